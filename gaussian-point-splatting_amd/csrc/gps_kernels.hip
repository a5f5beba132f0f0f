// gps_kernels.hip — gfx950 kernels of the per-frame hot path.
//
//   clear_kernel        FB <- background word          (gaussian_point_splatting.cu:580-593)
//   cull_kernel         block frustum + Hi-Z test      (occlusion/gaussian_bvh.cu:287-329)
//   preprocess_kernel   project, count, colour, cache  (gaussian_point_splatting.cu:147-307)
//   tile_scan_kernel +  workload distribution          (workload/workload_distributor.cu:62-145)
//   distribute_kernel   -> compacted (id, offset) lists + per-chunk start
//   splat_kernel        point sampling + 64-bit atomicMin (gaussian_point_splatting.cu:67-141,383-432)
//   combine_kernel      FB -> colour / NDC depth       (gaussian_point_splatting.cu:485-548)
//   hiz_build_kernel +  Hi-Z pyramid in two launches   (occlusion/depth_mip_chain.cu:107-145)
//   hiz_top_kernel
//   taa / progressive   resolve                        (passes/post_processing_kernels.cuh:108-285)
//
// Wave64: every ballot yields two 32-bit mask words so the reference's
// 32-flags-per-word layout (config.h:28) is kept.
#include "gps_fb.h"
#include "gps_internal.h"
#include "gps_kernels.h"
#include "gps_rng.h"

namespace gps {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ uint32_t wave_reduce_add(uint32_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Adds a per-thread count to *counter with one atomic per workgroup (the
// counter is a single word: per-wave atomics on it serialize).
__device__ __forceinline__ void add_count(uint32_t* counter, uint32_t v) {
    __shared__ uint32_t s_tot;
    if (threadIdx.x == 0) s_tot = 0u;
    __syncthreads();
    const uint32_t w = wave_reduce_add(v);
    if ((threadIdx.x & 63) == 0 && w) atomicAdd(&s_tot, w);
    __syncthreads();
    if (threadIdx.x == 0 && s_tot) atomicAdd(counter, s_tot);
}

__device__ __forceinline__ bool mask_bit(const uint32_t* m, uint32_t i) { return ((m[i >> 5] >> (i & 31u)) & 1u) != 0u; }

// Block culling of one preprocess workgroup's Gaussians [g0, g0 + threads):
// Gaussian g lies in block g / gpb, whose bit is bit b % 32 of mask word
// b / 32 (the same bit as block_culled below).  The workgroup's first block
// and its offset into it are found once (scalar division); a lane's block
// is then b0 + (r0 + lane) / gpb, a quotient of two integers below 2^24
// (exact in float: the quotient is at least 1 / gpb from the next integer,
// more than half an ulp) or, for blocks of 2^23 or more, 0 or 1.
struct WgBlocks {
    uint32_t b0, r0;
    bool some_culled;  // a block of the workgroup is culled (else no lane needs its own lookup)
};
__device__ __forceinline__ bool wg_all_culled(const CullView& c, uint32_t g0, uint32_t n_in_wg, WgBlocks& wb) {
    wb.b0 = g0 / c.gpb;
    wb.r0 = g0 - wb.b0 * c.gpb;
    const uint32_t b1 = (g0 + n_in_wg - 1u) / c.gpb;
    bool all = true, some = false;
    for (uint32_t b = wb.b0; b <= b1; ++b) {
        const bool cb = mask_bit(c.mask, b);
        all = all && cb;
        some = some || cb;
    }
    wb.some_culled = some;
    return all;
}
__device__ __forceinline__ bool lane_block_culled(const CullView& c, const WgBlocks& wb, uint32_t lane_off) {
    const uint32_t o = wb.r0 + lane_off;
    const uint32_t q = c.gpb >= (1u << 23) ? (o >= c.gpb ? 1u : 0u) : (uint32_t)((float)o / (float)c.gpb);
    return mask_bit(c.mask, wb.b0 + q);
}

// gaussian_bvh.cuh:47-53
__device__ __forceinline__ bool block_culled(const CullView& c, uint32_t g) {
    const uint32_t span = c.gpb * 32u;
    const uint32_t tree = g / span;
    const uint32_t blk = (g % span) / c.gpb;
    return ((c.mask[tree] >> blk) & 1u) != 0u;
}

// ------------------------------------------------------------------ scene access
__device__ __forceinline__ float plane(const DevScene& s, int p, uint32_t g) {
    return s.planes[(uint64_t)p * s.stride + g];
}
__device__ __forceinline__ float half_bits_to_float(uint32_t h16) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)h16);
}
struct Geo {
    float mx, my, mz, qw, qx, qy, qz, sx, sy, sz, op;
};
template <int FORMAT> __device__ __forceinline__ Geo load_geo(const DevScene& s, uint32_t g);
template <> __device__ __forceinline__ Geo load_geo<GPS_FORMAT_FP32>(const DevScene& s, uint32_t g) {
    Geo o;
    o.mx = plane(s, 0, g);
    o.my = plane(s, 1, g);
    o.mz = plane(s, 2, g);
    o.qw = plane(s, 3, g);
    o.qx = plane(s, 4, g);
    o.qy = plane(s, 5, g);
    o.qz = plane(s, 6, g);
    o.sx = plane(s, 7, g);
    o.sy = plane(s, 8, g);
    o.sz = plane(s, 9, g);
    o.op = plane(s, 10, g);
    return o;
}
template <> __device__ __forceinline__ Geo load_geo<GPS_FORMAT_HALF_FP8>(const DevScene& s, uint32_t g) {
    Geo o;
    o.mx = s.means[g];
    o.my = s.means[s.stride + g];
    o.mz = s.means[2 * s.stride + g];
    const uint4 q = s.geo[g];
    o.qw = half_bits_to_float(q.x & 0xFFFFu);
    o.qx = half_bits_to_float(q.x >> 16);
    o.qy = half_bits_to_float(q.y & 0xFFFFu);
    o.qz = half_bits_to_float(q.y >> 16);
    o.sx = half_bits_to_float(q.z & 0xFFFFu);
    o.sy = half_bits_to_float(q.z >> 16);
    o.sz = half_bits_to_float(q.w & 0xFFFFu);
    o.op = half_bits_to_float(q.w >> 16);
    return o;
}

// The reference's compressed scene (gaussian_types.cuh:323-374, 507-509):
// position = mix(box min, box max, code / (2^bits - 1)); rotation =
// quat_iem(2 * code / 1023 - 1) with the host path's arithmetic (:224-230:
// 1 / sqrt, sin, cos), stored (.x .. .w) like every other rotation; scale =
// exp(mix(log-scale box, code / 1023)); opacity = byte / 255.  glm::mix(x,
// y, a) = x * (1 - a) + y * a.
__device__ __forceinline__ float glm_mix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
template <> __device__ __forceinline__ Geo load_geo<GPS_FORMAT_QUANTIZED>(const DevScene& s, uint32_t g) {
    const uint4 w = s.geo[g];
    const float* b = s.qbounds;  // min_pos[3] max_pos[3] min_log_scale[3] max_log_scale[3] ...
    const uint32_t px = w.x & 0x7FFFFFu;
    const uint32_t py = ((w.x >> 23) & 0x1FFu) | ((w.y & 0x1FFFu) << 9);
    const uint32_t pz = ((w.y >> 13) & 0x7FFFFu) | ((w.z & 0xFu) << 19);
    Geo o;
    o.mx = glm_mix(b[0], b[3], (float)px / 8388607.0f);
    o.my = glm_mix(b[1], b[4], (float)py / 4194303.0f);
    o.mz = glm_mix(b[2], b[5], (float)pz / 8388607.0f);
    const uint32_t rx = (w.z >> 4) & 0x3FFu, ry = (w.z >> 14) & 0x3FFu;
    const uint32_t rz = ((w.z >> 24) & 0xFFu) | ((w.w & 0x3u) << 8);
    const float vx = ((float)rx / 1023.0f) * 2.0f - 1.0f;
    const float vy = ((float)ry / 1023.0f) * 2.0f - 1.0f;
    const float vz = ((float)rz / 1023.0f) * 2.0f - 1.0f;
    const float d = (vx * vx + vy * vy) + vz * vz;
    const float t = 1.0f / det_sqrt(d + 1e-8f);
    const float a = (1.57079632679f * t) * d;
    float sn, cs;
    det_sincosf_inrange(a, &sn, &cs);  // a in [0, 2.73]
    const float k = sn * t;
    o.qw = k * vx;
    o.qx = k * vy;
    o.qy = k * vz;
    o.qz = cs;
    o.sx = det_expf(glm_mix(b[6], b[9], (float)((w.w >> 2) & 0x3FFu) / 1023.0f));
    o.sy = det_expf(glm_mix(b[7], b[10], (float)((w.w >> 12) & 0x3FFu) / 1023.0f));
    o.sz = det_expf(glm_mix(b[8], b[11], (float)((w.w >> 22) & 0x3FFu) / 1023.0f));
    o.op = (float)s.qop[g] / 255.0f;
    return o;
}

// Sloan's real SH basis to degree 3 at a unit direction
// (eval_sh_sloan_l3, src/core/math/sh.cuh:55-96), same operation order.
__device__ __forceinline__ void sh_sloan_l3(float x, float y, float z, float o[16]) {
    const float z2 = z * z;
    o[0] = 0.282094792f;
    o[2] = 0.488602512f * z;
    float a = z2 - 0.333333333f;
    o[6] = 0.946174696f * a;
    const float b = z * (a - 0.266666667f);
    o[12] = 1.865881663f * b;
    float c1 = x, s1 = y;
    float d = -0.488602512f;
    o[1] = s1 * d;
    o[3] = c1 * d;
    d = -1.092548431f * z;
    o[5] = s1 * d;
    o[7] = c1 * d;
    a = z2 - 0.2f;
    d = -2.285228997f * a;
    o[11] = s1 * d;
    o[13] = c1 * d;
    const float c0 = x * c1 - y * s1;
    const float s0 = y * c1 + x * s1;
    d = 0.546274215f;
    o[4] = s0 * d;
    o[8] = c0 * d;
    d = 1.445305721f * z;
    o[10] = s0 * d;
    o[14] = c0 * d;
    c1 = x * c0 - y * s0;
    s1 = y * c0 + x * s0;
    d = -0.59004359f;
    o[9] = s1 * d;
    o[15] = c1 * d;
}

// SH colour of Gaussian g.
// Compact formats (kRecord) read one 64-byte SH record: load() requests it,
// eval_words() evaluates it, so a caller can issue the load early.
template <int FORMAT, int DEG> struct ShEval;
template <int DEG> struct ShEval<GPS_FORMAT_FP32, DEG> {
    static constexpr bool kRecord = false;
    __device__ __forceinline__ static void eval(const DevScene& s, uint32_t g, float x, float y, float z, float rgb[3]) {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch)
            rgb[ch] = sh_channel<DEG>(x, y, z, [&](int i) { return plane(s, 11 + 3 * i + ch, g); });
    }
};
template <int DEG> struct ShEval<GPS_FORMAT_HALF_FP8, DEG> {
    static constexpr bool kRecord = true;
    __device__ __forceinline__ static void load(const DevScene& s, uint32_t g, uint32_t words[16]) {
        const uint4* r = s.shrec + 4 * (uint64_t)g;
        const uint4 w0 = r[0];
        words[0] = w0.x; words[1] = w0.y; words[2] = w0.z; words[3] = w0.w;
        if (DEG > 0) {
            const uint4 w1 = r[1];
            words[4] = w1.x; words[5] = w1.y; words[6] = w1.z; words[7] = w1.w;
        }
        if (DEG > 1) {
            const uint4 w2 = r[2];
            words[8] = w2.x; words[9] = w2.y; words[10] = w2.z; words[11] = w2.w;
        }
        if (DEG > 2) {
            const uint4 w3 = r[3];
            words[12] = w3.x; words[13] = w3.y; words[14] = w3.z; words[15] = w3.w;
        }
    }
    __device__ __forceinline__ static void eval(const DevScene& s, uint32_t g, float x, float y, float z, float rgb[3]) {
        uint32_t words[16];
        load(s, g, words);
        eval_words(s, words, x, y, z, rgb);
    }
    __device__ __forceinline__ static void eval_words(const DevScene&, const uint32_t words[16], float x, float y, float z,
                                                      float rgb[3]) {
        const float scale = __builtin_bit_cast(float, words[2]);
        const float dc[3] = {half_bits_to_float(words[0] & 0xFFFFu), half_bits_to_float(words[0] >> 16),
                             half_bits_to_float(words[1] & 0xFFFFu)};
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            rgb[ch] = sh_channel<DEG>(x, y, z, [&](int i) -> float {
                if (i == 0) return dc[ch];
                const int b = 3 * (i - 1) + ch;           // byte index into the rest block
                const uint32_t w = words[3 + (b >> 2)];    // rest bytes start at byte 12
                return fp8_e4m3_to_float_hw(w, b & 3) * scale;
            });
        }
    }
};

// computeColorFromCompactSH (sh.cuh:257-273): all 16 coefficients whatever
// the scene's degree, summed in order from 0; coefficient i > 0 of channel c
// is mix(min_sh, max_sh, byte * (1/255)) (CompactSH::get, sh.cuh:41-47).
template <int DEG> struct ShEval<GPS_FORMAT_QUANTIZED, DEG> {
    static constexpr bool kRecord = true;
    __device__ __forceinline__ static void load(const DevScene& s, uint32_t g, uint32_t words[16]) {
        const uint4* r = s.shrec + 4 * (uint64_t)g;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 v = r[q];
            words[4 * q] = v.x;
            words[4 * q + 1] = v.y;
            words[4 * q + 2] = v.z;
            words[4 * q + 3] = v.w;
        }
    }
    __device__ __forceinline__ static void eval(const DevScene& s, uint32_t g, float x, float y, float z, float rgb[3]) {
        uint32_t words[16];
        load(s, g, words);
        eval_words(s, words, x, y, z, rgb);
    }
    __device__ __forceinline__ static void eval_words(const DevScene& s, const uint32_t words[16], float x, float y,
                                                      float z, float rgb[3]) {
        float e[16];
        sh_sloan_l3(x, y, z, e);
        const float* mn = s.qbounds + 12;
        const float* mx = s.qbounds + 57;
        const float inv255 = 1.0f / 255.0f;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            float acc = 0.0f;
            acc = acc + e[0] * __builtin_bit_cast(float, words[ch]);
#pragma unroll
            for (int i = 1; i < 16; ++i) {
                const int bi = 3 * (i - 1) + ch;  // code byte, after the 12 DC bytes
                const uint32_t code = (words[3 + (bi >> 2)] >> (8 * (bi & 3))) & 0xFFu;
                acc = acc + e[i] * glm_mix(mn[bi], mx[bi], (float)code * inv255);
            }
            acc = acc + 0.5f;
            rgb[ch] = (acc < 0.0f) ? 0.0f : acc;
        }
    }
};

// ------------------------------------------------------------------ clear
// Also zeroes up to 256 words at z0 and at z1 (the pass counters) in
// workgroup 0, which saves two one-workgroup launches per frame.
__global__ __launch_bounds__(256) void clear_kernel(uint4* __restrict__ fb, uint64_t n2, uint64_t value,
                                                    uint32_t* __restrict__ z0, uint32_t n0, uint32_t* __restrict__ z1,
                                                    uint32_t n1) {
    if (blockIdx.x == 0) {
        if (threadIdx.x < n0) z0[threadIdx.x] = 0u;
        if (threadIdx.x < n1) z1[threadIdx.x] = 0u;
    }
    const uint4 v = make_uint4((uint32_t)value, (uint32_t)(value >> 32), (uint32_t)value, (uint32_t)(value >> 32));
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (uint64_t)gridDim.x * blockDim.x)
        fb[i] = v;
}

// ------------------------------------------------------------------ block cull
struct Planes {
    float p[24];
};
// Appends the lanes with `pred` to a list with one atomic per workgroup on
// the list's counter; returns each such lane's slot.
__device__ __forceinline__ uint32_t wg_append(uint32_t* counter, bool pred) {
    __shared__ uint32_t s_wave[4], s_base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t b = __ballot(pred);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            const uint32_t c = s_wave[w];
            s_wave[w] = t;
            t += c;
        }
        s_base = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    return s_base + s_wave[wave] + __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
}

// `live` (or null): the unculled blocks are listed there, their count in
// *live_count (the group occlusion test runs over their groups only).
// `mz.wg_live` (or null): for a preprocess whose workgroups lie in single
// blocks (gpb a multiple of its 256 Gaussians), the cull writes each
// workgroup's live byte (0: its block is culled) and zeroes the valid-mask
// words of the culled blocks in the preprocess's current mask (maskA / maskB
// by the ping-pong index).
struct MaskZero {
    uint32_t* maskA;
    uint32_t* maskB;
    const uint32_t* iteration;
    uint8_t* wg_live;
    uint32_t N;
};
__global__ __launch_bounds__(256) void cull_kernel(CullView cv, Planes pl, Mat4 vp, HiZView h, int occl,
                                                   uint32_t* __restrict__ culled_count, uint32_t* __restrict__ live,
                                                   uint32_t* __restrict__ live_count, MaskZero mz) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    bool culled = false;
    if (idx < cv.node_count) {
        const CullNode n = cv.nodes[idx];
        bool outside = false;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const float nx = pl.p[4 * i], ny = pl.p[4 * i + 1], nz = pl.p[4 * i + 2], d = pl.p[4 * i + 3];
            const float vx = nx >= 0.0f ? n.mx[0] : n.mn[0];
            const float vy = ny >= 0.0f ? n.mx[1] : n.mn[1];
            const float vz = nz >= 0.0f ? n.mx[2] : n.mn[2];
            if (((nx * vx + ny * vy) + nz * vz) + d < 0.0f) outside = true;
        }
        if (outside) {
            culled = true;
        } else if (occl) {
            // gaussian_bvh.cu:237-285
            int minx = h.width - 1, miny = h.height - 1, maxx = 0, maxy = 0;
            float minDepth = 1e30f;
            bool anyFront = false;
            const float* M = vp.m;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const float cx = (i & 1) ? n.mx[0] : n.mn[0];
                const float cy = (i & 2) ? n.mx[1] : n.mn[1];
                const float cz = (i & 4) ? n.mx[2] : n.mn[2];
                const float hx = (M[0] * cx + M[4] * cy) + (M[8] * cz + M[12] * 1.0f);
                const float hy = (M[1] * cx + M[5] * cy) + (M[9] * cz + M[13] * 1.0f);
                const float hz = (M[2] * cx + M[6] * cy) + (M[10] * cz + M[14] * 1.0f);
                const float hw = (M[3] * cx + M[7] * cy) + (M[11] * cz + M[15] * 1.0f);
                if (hw <= 0.0f) continue;
                anyFront = true;
                const float nxp = hx / hw, nyp = hy / hw, nzp = hz / hw;
                const int px = det_f2i_rz((nxp * 0.5f + 0.5f) * (float)h.width);
                const int py = det_f2i_rz((nyp * 0.5f + 0.5f) * (float)h.height);
                minx = min(minx, px);
                miny = min(miny, py);
                maxx = max(maxx, px);
                maxy = max(maxy, py);
                minDepth = det_fmin(minDepth, nzp);
            }
            if (anyFront) {
                minx = max(minx, 0);
                miny = max(miny, 0);
                maxx = min(maxx, h.width - 1);
                maxy = min(maxy, h.height - 1);
                if (!(minx > maxx || miny > maxy)) culled = hiz_occluded(h, minDepth, minx, miny, maxx, maxy);
            }
        }
    }
    const uint64_t b = __ballot(culled);
    const int lane = threadIdx.x & 63;
    if (lane == 0) {
        const uint32_t w0 = (idx - lane) >> 5;
        if (w0 < cv.mask_words) cv.mask[w0] = (uint32_t)b;
        if (w0 + 1 < cv.mask_words) cv.mask[w0 + 1] = (uint32_t)(b >> 32);
    }
    add_count(culled_count, culled ? 1u : 0u);  // one atomic per workgroup on the single counter
    if (live) {
        const bool keep = idx < cv.node_count && !culled;
        const uint32_t k = wg_append(live_count, keep);
        if (keep) live[k] = idx;
    }
    if (mz.wg_live) {
        if (idx < cv.node_count) {
            const uint32_t wpb = cv.gpb / 256u, wg0 = idx * wpb;  // the block's preprocess workgroups
            const uint32_t nwg = min(wg0 + wpb, (mz.N + 255u) / 256u) - wg0;
            for (uint32_t w = 0; w < nwg; ++w) mz.wg_live[wg0 + w] = culled ? 0u : 1u;
        }
        // The wave's 64 blocks own consecutive mask words (gpb / 32 each): the
        // lanes zero the culled blocks' words in consecutive 16-byte pieces.
        uint32_t* m = (*mz.iteration & 1u) ? mz.maskB : mz.maskA;
        const uint32_t q = cv.gpb >> 7;  // 16-byte pieces per block
        const uint32_t blk0 = idx - (uint32_t)lane, words = (mz.N + 31u) >> 5;
        for (uint32_t c = (uint32_t)lane; c < 64u * q; c += 64u) {
            const uint32_t bl = c / q, w = (blk0 + bl) * (cv.gpb >> 5) + (c - bl * q) * 4u;
            if (((b >> bl) & 1ull) && blk0 + bl < cv.node_count) {
                if (w + 4u <= words)
                    *reinterpret_cast<uint4*>(m + w) = make_uint4(0u, 0u, 0u, 0u);
                else
                    for (uint32_t j = w; j < words; ++j) m[j] = 0u;
            }
        }
    }
}

// ------------------------------------------------------------------ group occlusion test
// group_occluded (gps_hiz.h) per group of GPS_GROUP Gaussians, before the
// preprocess of each pass; the preprocess skips the marked groups.

// One wave per group: the box of the centres and the covariance bounds of the
// members, decoded exactly as the preprocess decodes them.
template <int FORMAT>
__global__ __launch_bounds__(256) void group_box_kernel(DevScene s, GroupBox* __restrict__ boxes, uint32_t count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t gi = t / GPS_GROUP;
    if (gi >= count) return;  // whole waves (count * 64 >= N)
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    float lam = 0.0f, sig[3] = {0.0f, 0.0f, 0.0f};
    bool ok = true;
    if (t < s.N) {
        const Geo ge = load_geo<FORMAT>(s, t);
        const float a[10] = {ge.mx, ge.my, ge.mz, ge.qw, ge.qx, ge.qy, ge.qz, ge.sx, ge.sy, ge.sz};
        ok = group_member_bounds(a, &lam, sig);
        lo[0] = hi[0] = ge.mx;
        lo[1] = hi[1] = ge.my;
        lo[2] = hi[2] = ge.mz;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            lo[i] = fminf(lo[i], __shfl_xor(lo[i], o, 64));
            hi[i] = fmaxf(hi[i], __shfl_xor(hi[i], o, 64));
            sig[i] = fmaxf(sig[i], __shfl_xor(sig[i], o, 64));
        }
        lam = fmaxf(lam, __shfl_xor(lam, o, 64));
    }
    ok = __ballot(!ok) == 0ull;
    if ((threadIdx.x & 63) == 0) {
        GroupBox b;
        for (int i = 0; i < 3; ++i) {
            b.lo[i] = lo[i];
            b.hi[i] = hi[i];
            b.sig[i] = sig[i];
        }
        b.lam = lam;
        b.ok = ok ? 1u : 0u;
        b.pad = 0.0f;
        boxes[gi] = b;
    }
}

// Without block culling: one thread per group.
__global__ __launch_bounds__(256) void group_cull_kernel(GroupView gv, FrameConsts f, HiZView h,
                                                         uint32_t* __restrict__ culled_count) {
    const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    bool occ = false;
    if (gi < gv.count) {
        occ = group_occluded(gv.boxes[gi], f, h);
        gv.occluded[gi] = occ ? 1u : 0u;
    }
    if (culled_count) add_count(culled_count, occ ? 1u : 0u);
}

// With block culling: the groups of the unculled blocks the block cull
// listed (grid-stride over list entries x groups per block).  A group that
// straddles two unculled blocks is written twice with the same value; it is
// counted with the block holding its first Gaussian.  The count (diagnostic,
// `culled_count` null when off) costs one atomic on one word per workgroup.
__global__ __launch_bounds__(256) void group_cull_live_kernel(GroupView gv, CullView cv, const uint32_t* __restrict__ live,
                                                              const uint32_t* __restrict__ live_count, uint32_t per_block,
                                                              uint32_t N, FrameConsts f, HiZView h,
                                                              uint32_t* __restrict__ culled_count) {
    const uint64_t items = (uint64_t)(*live_count) * per_block;
    uint32_t counted = 0;
    for (uint64_t it = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; it < items;
         it += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = live[it / per_block];
        const uint32_t first = b * cv.gpb;                                // block's first Gaussian
        const uint32_t last = min(N, first + cv.gpb) - 1u;                // and last
        const uint32_t gi = first / GPS_GROUP + (uint32_t)(it % per_block);
        if (gi > last / GPS_GROUP) continue;
        const bool occ = group_occluded(gv.boxes[gi], f, h);
        gv.occluded[gi] = occ ? 1u : 0u;
        counted += (occ && gi * GPS_GROUP >= first) ? 1u : 0u;
    }
    if (culled_count) add_count(culled_count, counted);
}

// ------------------------------------------------------------------ preprocess
// One Gaussian under one frame (gaussian_point_splatting.cu:147-307, steps 2-11
// of its cull order): returns the point count, 0 if the Gaussian is not
// splatted, and writes its 64-byte splat record when the count is > 0.
// `mask_prev` is read only when f.only_prev is set; `h` only when f.occl.
template <int FORMAT, int DEG>
__device__ __forceinline__ uint32_t preprocess_one(const DevScene& s, const FrameConsts& f, uint32_t g,
                                                   const uint32_t* __restrict__ mask_prev, const HiZView& h,
                                                   SplatRecord* __restrict__ rec_out) {
    if (f.only_prev && mask_bit(mask_prev, g)) return 0u;
    u32x2 seed = make_seed(g, f.frame_seed);
    const Geo ge = load_geo<FORMAT>(s, g);
    const float op = glm_min(ge.op, 1.0f);
    if (op <= 0.0f) return 0u;
    Projected pr;
    if (!project_gaussian(f, ge.mx, ge.my, ge.mz, ge.qw, ge.qx, ge.qy, ge.qz, ge.sx, ge.sy, ge.sz, pr)) return 0u;
    if (pr.minx >= f.up_w || pr.miny >= f.up_h || pr.maxx < 0 || pr.maxy < 0) return 0u;
    if (f.cull_small) {
        const float tr = pr.ca + pr.cc;
        const float t2 = 0.5f * tr;
        const float term2 = det_sqrt(det_fmax(0.1f, t2 * t2 - pr.det));
        const float e1 = t2 + term2, e2 = t2 - term2;
        if (e1 < 0.02f || e2 < 0.02f) return 0u;
    }
    // (The Hi-Z test is a pure rejection, so it is taken before the
    // importance terms it does not depend on.)
    if (f.occl && hiz_occluded(h, pr.pz, div_ss(f, pr.minx), div_ss(f, pr.miny), div_ss(f, pr.maxx), div_ss(f, pr.maxy))) return 0u;
    const float s_det = det_sqrt(pr.det);
    const float dl = dilog(op);
    const float importance = f.unbiased ? GPS_TWO_PI_F * s_det * dl : op * GPS_TWO_PI_F * s_det;
    uint32_t cnt = f.unbiased ? poisson(seed, importance) : (uint32_t)stochastic_round(importance, pcg2d(seed).x);
    if (f.K > 1.0f) cnt = (uint32_t)stochastic_round((float)cnt / f.K, pcg2d(seed).x);
    if (cnt == 0u) return 0u;
    cnt = min(cnt, f.point_cap);
    float dx, dy, dz;
    view_dir(f, ge.mx, ge.my, ge.mz, dx, dy, dz);
    float rgb[3];
    ShEval<FORMAT, DEG>::eval(s, g, dx, dy, dz, rgb);
    const float c0 = det_sqrt(pr.ca);
    const float c1 = pr.cb / c0;
    const float c2 = det_sqrt(pr.cc - c1 * c1);
    const float inv_det = 1.0f / pr.det;
    const uint64_t packed = (depth_quantize(pr.view_z, f.near_plane, f.far_plane) << GPS_DEPTH_SHIFT) |
                            pack_color36(rgb[0], rgb[1], rgb[2]);
    uint4* r = reinterpret_cast<uint4*>(rec_out);
    r[0] = make_uint4(__float_as_uint(pr.pmx), __float_as_uint(pr.pmy), __float_as_uint(c0), __float_as_uint(c1));
    r[1] = make_uint4(__float_as_uint(c2), __float_as_uint(pr.cc * inv_det), __float_as_uint(-pr.cb * inv_det),
                      __float_as_uint(pr.ca * inv_det));
    r[2] = make_uint4(__float_as_uint(op), __float_as_uint(corrected_inv_alpha(op)), __float_as_uint(dl),
                      (uint32_t)pr.minx | ((uint32_t)pr.miny << 16));
    r[3] = make_uint4((uint32_t)pr.maxx | ((uint32_t)pr.maxy << 16), (uint32_t)packed, (uint32_t)(packed >> 32),
                      __float_as_uint(alpha_threshold(op)));
    return cnt;
}

// The single-view preprocess runs the steps of preprocess_one in two phases
// with the Gaussians that survive the first repacked onto the first lanes of
// the workgroup (an LDS work list), so the counting and the colour steps run
// on full waves instead of the few lanes that need them:
//   A  every lane: block cull, previous-mask skip, opacity, projection,
//      bounds, small-eigenvalue cull, Hi-Z test -> list 1
//   B  list 1: importance, Poisson count, / K, cap; weight; for counts > 0
//      the record and its packed depth | colour word (SH colour).
// Per Gaussian the arithmetic is exactly preprocess_one's.
struct PreItem {
    uint32_t g;
    float op, view_z, pmx, pmy, ca, cb, cc, det;
    uint32_t bmin, bmax;
    float mx, my, mz;  // world mean (the colour's view direction)
};
__device__ __forceinline__ void push_item(PreItem* s_item, uint32_t* s_n, uint32_t g, float op, const Projected& pr,
                                          float mx, float my, float mz) {
    PreItem& q = s_item[atomicAdd(s_n, 1u)];
    q.g = g;
    q.op = op;
    q.view_z = pr.view_z;
    q.pmx = pr.pmx;
    q.pmy = pr.pmy;
    q.ca = pr.ca;
    q.cb = pr.cb;
    q.cc = pr.cc;
    q.det = pr.det;
    q.bmin = (uint32_t)pr.minx | ((uint32_t)pr.miny << 16);
    q.bmax = (uint32_t)pr.maxx | ((uint32_t)pr.maxy << 16);
    q.mx = mx;
    q.my = my;
    q.mz = mz;
}

// Two 32-bit valid-mask words per wave and the workgroup's (points, valid)
// sums added to its 4096-Gaussian tile; n = this lane's point count.
__device__ __forceinline__ void preprocess_epilogue(uint32_t n, uint32_t g, uint32_t g0, uint32_t N,
                                                    uint32_t* __restrict__ mask_cur, uint32_t* __restrict__ tile_sum,
                                                    uint32_t* __restrict__ tile_cnt) {
    const int tid = threadIdx.x;
    const uint64_t b = __ballot(n != 0u);
    const int lane = tid & 63;
    const uint32_t wave_first = g - (uint32_t)lane;
    const uint32_t sum = wave_reduce_add(n);
    __shared__ uint32_t s_sum[GPS_PRE_THREADS / 64], s_cnt[GPS_PRE_THREADS / 64];
    if (lane == 0) {
        if (wave_first < N) mask_cur[wave_first >> 5] = (uint32_t)b;
        if (wave_first + 32u < N) mask_cur[(wave_first >> 5) + 1] = (uint32_t)(b >> 32);
        s_sum[tid >> 6] = sum;
        s_cnt[tid >> 6] = (uint32_t)__popcll(b);
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t ts = 0, tc = 0;
#pragma unroll
        for (int w = 0; w < GPS_PRE_THREADS / 64; ++w) {
            ts += s_sum[w];
            tc += s_cnt[w];
        }
        const uint32_t tile = g0 >> GPS_TILE_SHIFT;
        if (tc) {
            atomicAdd(tile_sum + tile, ts);
            atomicAdd(tile_cnt + tile, tc);
        }
    }
}

// Reject-reason histogram of a pass (diagnostic, STATS = true): why each
// Gaussian of an unculled block did not become valid, in the cull order of
// gaussian_point_splatting.cu:188-263.  One counter set per pass.
enum RejectReason : uint32_t {
    REJ_TESTED = 0,     // Gaussians of unculled blocks (every lane below counts here too)
    REJ_PREV_VALID = 1, // second pass: valid in the first pass's mask
    REJ_OPACITY = 2,    // opacity <= 0
    REJ_PROJECT = 3,    // near plane, NDC z outside [-1, 1], det <= 0
    REJ_BOUNDS = 4,     // footprint outside the screen
    REJ_SMALL = 5,      // small-eigenvalue cull
    REJ_HIZ = 6,        // occluded by the Hi-Z chain
    REJ_ZERO = 7,       // point count 0
    REJ_VALID = 8,      // valid (count > 0)
    REJ_COUNT = 9
};
// Adds one per lane whose `reason` is r to counter r (r = REJ_TESTED adds every
// lane with a reason); one LDS atomic per wave and reason, one global atomic
// per workgroup and reason.
__device__ __forceinline__ void count_reasons(uint32_t* s_rej, uint32_t reason, bool tested) {
    const int lane = threadIdx.x & 63;
    const uint32_t nt = (uint32_t)__popcll(__ballot(tested));
    if (lane == 0 && nt) atomicAdd(s_rej + REJ_TESTED, nt);
#pragma unroll
    for (uint32_t r = 1; r < REJ_COUNT; ++r) {
        const uint32_t n = (uint32_t)__popcll(__ballot(reason == r));
        if (lane == 0 && n) atomicAdd(s_rej + r, n);
    }
}

// The preprocess of the workgroup's Gaussians [g0, g0 + GPS_PRE_THREADS).
// `live`: they lie in a block the block cull kept (no block lookup).
template <int FORMAT, int DEG, bool STATS>
__device__ __forceinline__ void preprocess_wg(const DevScene& s, const FrameConsts& f, uint32_t* __restrict__ mask_cur,
                                              const uint32_t* __restrict__ mask_prev, uint32_t* __restrict__ weights,
                                              SplatRecord* __restrict__ recs, uint32_t* __restrict__ tile_sum,
                                              uint32_t* __restrict__ tile_cnt, const CullView& cv, const HiZView& h,
                                              const uint8_t* __restrict__ gocc_bytes, uint32_t* __restrict__ rejects,
                                              uint32_t g0, bool live) {
    __shared__ PreItem s_item[GPS_PRE_THREADS];
    // 15 KB of LDS per workgroup
    __shared__ uint32_t s_count[GPS_PRE_THREADS];
    __shared__ uint32_t s_n1;
    const int tid = threadIdx.x;
    const uint32_t g = g0 + (uint32_t)tid;
    WgBlocks wb{0u, 0u, false};
    if (!live && f.hier && wg_all_culled(cv, g0, min((uint32_t)GPS_PRE_THREADS, s.N - g0), wb)) {
        // Every Gaussian here is culled: only the valid-mask words (zero)
        // are written; weights stay untouched, the tile sums get nothing.
        if (tid < GPS_PRE_THREADS / 32 && g0 + 32u * (uint32_t)tid < s.N) mask_cur[(g0 >> 5) + (uint32_t)tid] = 0u;
        return;
    }
    // Occluded groups (group_cull_kernel; one group per wave): their members
    // are rejected by the Hi-Z test or earlier, so they take weight 0 here.
    // The counting form runs every step and checks that claim instead.
    static_assert(GPS_PRE_THREADS / GPS_GROUP == 4 && GPS_GROUP == 64, "a workgroup's four group bytes in one word");
    constexpr uint32_t kAllGroups = 0xFu;
    uint32_t gbits = 0u;
    if (gocc_bytes) {
        const uint32_t w = reinterpret_cast<const uint32_t*>(gocc_bytes)[g0 / GPS_PRE_THREADS];
        gbits = (w & 1u) | ((w >> 7) & 2u) | ((w >> 14) & 4u) | ((w >> 21) & 8u);
    }
    const bool lane_blk = f.hier && wb.some_culled && lane_block_culled(cv, wb, (uint32_t)tid);
    const bool gocc = ((gbits >> (tid / GPS_GROUP)) & 1u) != 0u;
    if (!STATS && gbits == kAllGroups) {
        if (g < s.N && !lane_blk) weights[g] = 0u;
        if (tid < GPS_PRE_THREADS / 32 && g0 + 32u * (uint32_t)tid < s.N) mask_cur[(g0 >> 5) + (uint32_t)tid] = 0u;
        return;
    }
    __shared__ uint32_t s_rej[STATS ? REJ_COUNT + 1 : 1];
    if (tid == 0) {
        s_n1 = 0u;
    }
    if (STATS && tid <= REJ_COUNT) s_rej[tid] = 0u;
    s_count[tid] = 0u;
    __syncthreads();
    // Phase A.  Gaussians of culled blocks leave their weight slot untouched:
    // the distributor reads the same cull mask and counts them as zero.
    uint32_t reason = REJ_TESTED;  // STATS: this lane's reject reason (REJ_TESTED = none yet)
    const bool tested = g < s.N && !lane_blk && (STATS || !gocc);
    if (!STATS && gocc && g < s.N && !lane_blk) weights[g] = 0u;
    if constexpr (STATS) {
    if (tested) {
        bool survives = false;
        if (f.only_prev && mask_bit(mask_prev, g)) {
            reason = REJ_PREV_VALID;
        } else {
            const Geo ge = load_geo<FORMAT>(s, g);
            const float op = glm_min(ge.op, 1.0f);
            Projected pr;
            bool projected = false;
            if (!(op > 0.0f)) {
                reason = REJ_OPACITY;
            } else if (!(projected = project_gaussian(f, ge.mx, ge.my, ge.mz, ge.qw, ge.qx, ge.qy, ge.qz, ge.sx,
                                                      ge.sy, ge.sz, pr))) {
                reason = REJ_PROJECT;
            } else if (pr.minx >= f.up_w || pr.miny >= f.up_h || pr.maxx < 0 || pr.maxy < 0) {
                reason = REJ_BOUNDS;
            }
            if (projected && reason == REJ_TESTED) {
                survives = true;
                if (f.cull_small) {
                    const float tr = pr.ca + pr.cc;
                    const float t2 = 0.5f * tr;
                    const float term2 = det_sqrt(det_fmax(0.1f, t2 * t2 - pr.det));
                    const float e1 = t2 + term2, e2 = t2 - term2;
                    if (e1 < 0.02f || e2 < 0.02f) {
                        survives = false;
                        reason = REJ_SMALL;
                    }
                }
                if (survives && f.occl &&
                    hiz_occluded(h, pr.pz, div_ss(f, pr.minx), div_ss(f, pr.miny), div_ss(f, pr.maxx), div_ss(f, pr.maxy))) {
                    survives = false;
                    reason = REJ_HIZ;
                }
                if (survives) push_item(s_item, &s_n1, g, op, pr, ge.mx, ge.my, ge.mz);
            }
        }
        if (!survives) weights[g] = 0u;
    }
    } else {
    // The same steps without the reasons.  The previous-mask word, the
    // attributes and the projection's centre are all needed before the first
    // branch, so their loads go out together (one memory round trip).
    if (tested) {
        bool survives = false;
        const uint32_t prev_word = mask_prev[g >> 5];  // read in both passes: no branch before the other loads
        const Geo ge = load_geo<FORMAT>(s, g);
        const bool prev = f.only_prev && ((prev_word >> (g & 31u)) & 1u) != 0u;
        const float op = glm_min(ge.op, 1.0f);
        Center ce;
        const bool front = project_center(f, ge.mx, ge.my, ge.mz, ce);
        Projected pr;
        if (!prev && op > 0.0f && front)
            survives = project_footprint(f, ce, ge.qw, ge.qx, ge.qy, ge.qz, ge.sx, ge.sy, ge.sz, pr) &&
                       !(pr.minx >= f.up_w || pr.miny >= f.up_h || pr.maxx < 0 || pr.maxy < 0);
        if (survives && f.cull_small) {
            const float tr = pr.ca + pr.cc;
            const float t2 = 0.5f * tr;
            const float term2 = det_sqrt(det_fmax(0.1f, t2 * t2 - pr.det));
            const float e1 = t2 + term2, e2 = t2 - term2;
            survives = !(e1 < 0.02f || e2 < 0.02f);
        }
        if (survives && f.occl)
            survives = !hiz_occluded(h, pr.pz, div_ss(f, pr.minx), div_ss(f, pr.miny), div_ss(f, pr.maxx),
                                     div_ss(f, pr.maxy));
        if (survives)
            push_item(s_item, &s_n1, g, op, pr, ge.mx, ge.my, ge.mz);
        else
            weights[g] = 0u;
    }
    }
    if (STATS) {
        count_reasons(s_rej, reason, tested);
        const uint32_t nv = (uint32_t)__popcll(__ballot(tested && gocc && reason == REJ_TESTED));
        if ((tid & 63) == 0 && nv) atomicAdd(s_rej + REJ_COUNT, nv);  // survivors of occluded groups
    }
    __syncthreads();
    // Phase B: importance, count, record and, for counts > 0, colour.  The
    // compact formats' 64-byte SH record is requested before the count, so
    // its latency overlaps the count's arithmetic.
    const bool in1 = (uint32_t)tid < s_n1;
    PreItem q;
    if (in1) q = s_item[tid];
    uint32_t cnt_b = 0u;  // this lane's count (STATS)
    if (in1) {
        uint32_t shw[16];
        if constexpr (ShEval<FORMAT, DEG>::kRecord) ShEval<FORMAT, DEG>::load(s, q.g, shw);
        u32x2 seed = make_seed(q.g, f.frame_seed);
        const float s_det = det_sqrt(q.det);
        const float dl = dilog(q.op);
        const float importance = f.unbiased ? GPS_TWO_PI_F * s_det * dl : q.op * GPS_TWO_PI_F * s_det;
        uint32_t cnt = f.unbiased ? poisson(seed, importance) : (uint32_t)stochastic_round(importance, pcg2d(seed).x);
        if (f.K > 1.0f) cnt = (uint32_t)stochastic_round((float)cnt / f.K, pcg2d(seed).x);
        if (cnt) {
            cnt = min(cnt, f.point_cap);
            const float c0 = det_sqrt(q.ca);
            const float c1 = q.cb / c0;
            const float c2 = det_sqrt(q.cc - c1 * c1);
            const float inv_det = 1.0f / q.det;
            uint4* r = reinterpret_cast<uint4*>(recs + q.g);
            r[0] = make_uint4(__float_as_uint(q.pmx), __float_as_uint(q.pmy), __float_as_uint(c0), __float_as_uint(c1));
            r[1] = make_uint4(__float_as_uint(c2), __float_as_uint(q.cc * inv_det), __float_as_uint(-q.cb * inv_det),
                              __float_as_uint(q.ca * inv_det));
            r[2] = make_uint4(__float_as_uint(q.op), __float_as_uint(corrected_inv_alpha(q.op)), __float_as_uint(dl),
                              q.bmin);
            float dx, dy, dz;
            view_dir(f, q.mx, q.my, q.mz, dx, dy, dz);
            float rgb[3];
            if constexpr (ShEval<FORMAT, DEG>::kRecord)
                ShEval<FORMAT, DEG>::eval_words(s, shw, dx, dy, dz, rgb);
            else
                ShEval<FORMAT, DEG>::eval(s, q.g, dx, dy, dz, rgb);
            const uint64_t packed = ((uint64_t)(uint32_t)depth_quantize(q.view_z, f.near_plane, f.far_plane)
                                     << GPS_DEPTH_SHIFT) | pack_color36(rgb[0], rgb[1], rgb[2]);
            r[3] = make_uint4(q.bmax, (uint32_t)packed, (uint32_t)(packed >> 32), __float_as_uint(alpha_threshold(q.op)));
            s_count[q.g - g0] = cnt;
        }
        weights[q.g] = cnt;
        cnt_b = cnt;
    }
    if (STATS) {
        const uint32_t nz = (uint32_t)__popcll(__ballot(in1 && cnt_b == 0u));
        const uint32_t nv = (uint32_t)__popcll(__ballot(in1 && cnt_b != 0u));
        if ((tid & 63) == 0) {
            if (nz) atomicAdd(s_rej + REJ_ZERO, nz);
            if (nv) atomicAdd(s_rej + REJ_VALID, nv);
        }
    }
    __syncthreads();
    if (STATS && tid < REJ_COUNT && s_rej[tid]) atomicAdd(rejects + (f.only_prev ? REJ_COUNT : 0) + tid, s_rej[tid]);
    if (STATS && tid == REJ_COUNT && s_rej[tid]) atomicAdd(rejects + 2 * REJ_COUNT + (f.only_prev ? 1 : 0), s_rej[tid]);
    preprocess_epilogue(s_count[tid], g, g0, s.N, mask_cur, tile_sum, tile_cnt);
}

// One workgroup per GPS_PRE_THREADS Gaussians.  `wg_live` (or null): one
// byte per workgroup written by the block cull, 0 when its block is culled;
// the cull also zeroed such a workgroup's valid-mask words, so it returns
// after that one load (most of the grid in culled frames).
template <int FORMAT, int DEG, bool STATS>
__global__ __launch_bounds__(GPS_PRE_THREADS) void preprocess_kernel(
    DevScene s, FrameConsts f, uint32_t* __restrict__ maskA, uint32_t* __restrict__ maskB,
    const uint32_t* __restrict__ iteration, uint32_t* __restrict__ weights, SplatRecord* __restrict__ recs,
    uint32_t* __restrict__ tile_sum, uint32_t* __restrict__ tile_cnt, CullView cv, HiZView h,
    const uint8_t* __restrict__ gocc_bytes, uint32_t* __restrict__ rejects, const uint8_t* __restrict__ wg_live) {
    if (wg_live && wg_live[blockIdx.x] == 0u) return;
    // workload_distributor.cu:32-40: current = A on even iterations.
    const uint32_t it = *iteration;
    uint32_t* mask_cur = (it & 1u) ? maskB : maskA;
    const uint32_t* mask_prev = (it & 1u) ? maskA : maskB;
    preprocess_wg<FORMAT, DEG, STATS>(s, f, mask_cur, mask_prev, weights, recs, tile_sum, tile_cnt, cv, h, gocc_bytes,
                                      rejects, blockIdx.x * GPS_PRE_THREADS, wg_live != nullptr);
}

// The same outputs with preprocess_one per lane and no work lists: cheaper
// when nearly every Gaussian survives projection (frames without block or
// occlusion culling, where the lists only add synchronisation).
template <int FORMAT, int DEG>
__global__ __launch_bounds__(GPS_PRE_THREADS) void preprocess_flat_kernel(
    DevScene s, FrameConsts f, uint32_t* __restrict__ maskA, uint32_t* __restrict__ maskB,
    const uint32_t* __restrict__ iteration, uint32_t* __restrict__ weights, SplatRecord* __restrict__ recs,
    uint32_t* __restrict__ tile_sum, uint32_t* __restrict__ tile_cnt, CullView cv, HiZView h) {
    const uint32_t g0 = blockIdx.x * GPS_PRE_THREADS;
    const uint32_t g = g0 + threadIdx.x;
    const uint32_t it = *iteration;
    uint32_t* mask_cur = (it & 1u) ? maskB : maskA;
    const uint32_t* mask_prev = (it & 1u) ? maskA : maskB;
    uint32_t n = 0;
    if (g < s.N && !(f.hier && block_culled(cv, g))) {
        n = preprocess_one<FORMAT, DEG>(s, f, g, mask_prev, h, recs + g);
        weights[g] = n;
    }
    preprocess_epilogue(n, g, g0, s.N, mask_cur, tile_sum, tile_cnt);
}

// ------------------------------------------------------------------ distribution
// Exclusive scan of the per-tile (point, valid) sums, one workgroup per step
// of 8192 tiles; the last workgroup publishes the pass totals.  Workgroup s
// takes its carry by reducing the raw sums of steps 0 .. s-1 itself, so no
// workgroup waits for another (sums are modulo 2^32, as a sequential scan's).
// The tile sums are cleared for the next pass by distribute_kernel, which
// runs after this kernel on every tile.
__global__ __launch_bounds__(1024) void tile_scan_kernel(const uint32_t* __restrict__ tile_sum,
                                                         const uint32_t* __restrict__ tile_cnt, uint32_t ntiles,
                                                         uint32_t* __restrict__ tile_off,
                                                         uint32_t* __restrict__ tile_voff, PassCounters* __restrict__ ctr,
                                                         uint32_t budget, uint32_t* __restrict__ iteration) {
    constexpr int PER = 8;  // tiles per thread
    // The step's 8192 tiles pass through LDS so that global loads and stores
    // are coalesced while each thread scans 8 consecutive tiles; one pad word
    // per 64 keeps the strided LDS reads conflict-free.  Loads take a clamped
    // index instead of a guard so that all of a thread's loads are in flight
    // together.
    constexpr uint32_t STEP = 1024u * PER;
    __shared__ uint32_t s_ta[STEP + STEP / 64], s_tc[STEP + STEP / 64];
    __shared__ uint32_t s_a[16], s_c[16], s_ra[16], s_rc[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    auto pidx = [](uint32_t j) { return j + (j >> 6); };
    const uint32_t base = blockIdx.x * STEP;  // base < ntiles unless ntiles == 0 (one workgroup)
    uint32_t va[PER], vc[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t i = base + (uint32_t)k * 1024u + (uint32_t)tid;
        const uint32_t ic = i < ntiles ? i : (ntiles ? ntiles - 1u : 0u);
        const uint32_t x = ntiles ? tile_sum[ic] : 0u, y = ntiles ? tile_cnt[ic] : 0u;
        va[k] = i < ntiles ? x : 0u;
        vc[k] = i < ntiles ? y : 0u;
    }
    // Carry: tiles [0, base) are whole steps.
    uint32_t ra = 0, rc = 0;
    for (uint32_t b = 0; b < base; b += STEP) {
        uint32_t xa[PER], xc[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            xa[k] = tile_sum[b + (uint32_t)k * 1024u + (uint32_t)tid];
            xc[k] = tile_cnt[b + (uint32_t)k * 1024u + (uint32_t)tid];
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            ra += xa[k];
            rc += xc[k];
        }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t j = (uint32_t)k * 1024u + (uint32_t)tid;
        s_ta[pidx(j)] = va[k];
        s_tc[pidx(j)] = vc[k];
    }
    ra = wave_reduce_add(ra);
    rc = wave_reduce_add(rc);
    if (lane == 0) {
        s_ra[wave] = ra;
        s_rc[wave] = rc;
    }
    __syncthreads();
    const uint32_t j0 = (uint32_t)tid * PER;
    uint32_t a[PER], c[PER], sa = 0, sc = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        a[k] = s_ta[pidx(j0 + (uint32_t)k)];
        c[k] = s_tc[pidx(j0 + (uint32_t)k)];
        sa += a[k];
        sc += c[k];
    }
    const uint32_t ia = wave_scan_incl(sa);
    const uint32_t ic = wave_scan_incl(sc);
    if (lane == 63) {
        s_a[wave] = ia;
        s_c[wave] = ic;
    }
    __syncthreads();
    uint32_t ca = 0, cc = 0, wa = 0, wc = 0;
    for (int w = 0; w < 16; ++w) {
        ca += s_ra[w];
        cc += s_rc[w];
    }
    for (int w = 0; w < wave; ++w) {
        wa += s_a[w];
        wc += s_c[w];
    }
    uint32_t oa = ca + wa + ia - sa, oc = cc + wc + ic - sc;
#pragma unroll
    for (int k = 0; k < PER; ++k) {  // exclusive prefixes back into LDS
        s_ta[pidx(j0 + (uint32_t)k)] = oa;
        s_tc[pidx(j0 + (uint32_t)k)] = oc;
        oa += a[k];
        oc += c[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t j = (uint32_t)k * 1024u + (uint32_t)tid;
        const uint32_t i = base + j;
        if (i < ntiles) {
            tile_off[i] = s_ta[pidx(j)];
            tile_voff[i] = s_tc[pidx(j)];
        }
    }
    if (blockIdx.x == gridDim.x - 1u && tid == 1023) {
        const uint32_t total = ca + wa + ia;
        const uint32_t P = total < budget ? total : budget;
        ctr->total = total;
        ctr->valid = cc + wc + ic;
        ctr->points = P;
        ctr->chunks = (uint32_t)(((uint64_t)P + GPS_SPLAT_CHUNK - 1) / GPS_SPLAT_CHUNK);
        // The reference flips the mask ping-pong only when the pass produced
        // points (workload_distributor.cu:111-115,144).
        if (P > 0u) *iteration = *iteration + 1u;
    }
}

// Per tile of 4096 Gaussians: local exclusive scan of the masked weights, then
// the compacted (Gaussian id, point offset) lists and the first compacted
// entry of every splat chunk.  Equivalent to the reference's scatter +
// inclusive max-scan: point p belongs to the last valid Gaussian whose offset
// is <= p.
// Tiles without valid Gaussians are skipped whole; with block culling on,
// weights of culled blocks are taken as zero without being read (preprocess
// does not write them).
//
// Several views can be concatenated (multi-view batches): view v owns tiles
// [v * tiles_per_view, (v + 1) * tiles_per_view), its Gaussian g sits at
// v * npad + g, its cull mask at cv.mask + v * mask_stride; `cview`, when not
// null, receives the view of every compacted entry.  Chunks are
// 2^chunk_shift points.
__global__ __launch_bounds__(256) void distribute_kernel(const uint32_t* __restrict__ weights,
                                                         const uint32_t* __restrict__ tile_off,
                                                         const uint32_t* __restrict__ tile_voff, uint32_t ntiles,
                                                         const PassCounters* __restrict__ ctr, uint32_t* __restrict__ cidx,
                                                         uint32_t* __restrict__ coff, uint32_t* __restrict__ chunk_start,
                                                         CullView cv, int hier, uint32_t N, uint32_t tiles_per_view,
                                                         uint32_t npad, uint32_t mask_stride,
                                                         uint8_t* __restrict__ cview, uint32_t chunk_shift,
                                                         uint32_t* __restrict__ tile_sum, uint32_t* __restrict__ tile_cnt) {
    __shared__ uint32_t s_a[4], s_c[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t tile = blockIdx.x;
    if (tid == 0) {  // the scan has consumed them: clear for the next pass
        tile_sum[tile] = 0u;
        tile_cnt[tile] = 0u;
    }
    // The workgroup's offsets and the pass totals are read up front, together.
    const uint32_t v0 = tile_voff[tile];
    const uint32_t vend = tile + 1 < ntiles ? tile_voff[tile + 1] : ctr->valid;
    const uint32_t a0 = tile_off[tile];
    const uint64_t P = ctr->points;
    if (vend == v0) return;  // uniform over the workgroup
    const uint32_t view = tile / tiles_per_view;
    const uint32_t* cmask = cv.mask + (uint64_t)view * mask_stride;
    const uint32_t g0 = (tile << GPS_TILE_SHIFT) + (uint32_t)tid * 16u;
    const uint32_t l0 = g0 - view * npad;  // index within the view's scene
    // bit k set: element g0 + k counts (inside the scene and not culled)
    uint32_t live = l0 >= N ? 0u : (N - l0 >= 16u ? 0xFFFFu : ((1u << (N - l0)) - 1u));
    if (hier && live) {
        const int kmax = N - l0 >= 16u ? 16 : (int)(N - l0);
        // The mask words of the thread's first and last block are read
        // together; blocks in between (fewer than 16 Gaussians per block)
        // are read when reached.
        const uint32_t bf = l0 / cv.gpb, bl = (l0 + (uint32_t)kmax - 1u) / cv.gpb;
        const uint32_t mf = cmask[bf >> 5], ml = cmask[bl >> 5];
        auto culled = [&](uint32_t b) {
            const uint32_t m = b == bf ? mf : (b == bl ? ml : cmask[b >> 5]);
            return ((m >> (b & 31u)) & 1u) != 0u;
        };
        uint32_t b = bf;
        uint32_t next = (b + 1u) * cv.gpb;
        bool c = culled(b);
        for (int k = 0; k < kmax; ++k) {
            if (l0 + (uint32_t)k == next) {
                ++b;
                next += cv.gpb;
                c = culled(b);
            }
            if (c) live &= ~(1u << k);
        }
    }
    const uint4* wp = reinterpret_cast<const uint4*>(weights + g0);
    // Groups of 4 without a live element re-read the first live group's
    // address (no extra cache line) so that the 4 loads need no branch and
    // are in flight together; their values are discarded.
    uint4 v4[4];
    if (live) {
        const int kl = __builtin_ctz(live) >> 2;
#pragma unroll
        for (int k = 0; k < 4; ++k) v4[k] = wp[((live >> (4 * k)) & 15u) ? k : kl];
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) v4[k] = make_uint4(0u, 0u, 0u, 0u);
    }
    uint32_t w[16];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint4 v = v4[k];
        w[4 * k] = (live >> (4 * k)) & 1u ? v.x : 0u;
        w[4 * k + 1] = (live >> (4 * k + 1)) & 1u ? v.y : 0u;
        w[4 * k + 2] = (live >> (4 * k + 2)) & 1u ? v.z : 0u;
        w[4 * k + 3] = (live >> (4 * k + 3)) & 1u ? v.w : 0u;
    }
    uint32_t sa = 0, sc = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        sa += w[k];
        sc += w[k] ? 1u : 0u;
    }
    const uint32_t ia = wave_scan_incl(sa);
    const uint32_t ic = wave_scan_incl(sc);
    if (lane == 63) {
        s_a[wave] = ia;
        s_c[wave] = ic;
    }
    __syncthreads();
    uint32_t wa = 0, wc = 0;
    for (int q = 0; q < wave; ++q) {
        wa += s_a[q];
        wc += s_c[q];
    }
    uint32_t off = a0 + wa + ia - sa;
    uint32_t r = v0 + wc + ic - sc;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (w[k]) {
            cidx[r] = g0 + (uint32_t)k;
            coff[r] = off;
            if (cview) cview[r] = (uint8_t)view;
            const uint64_t lo = off;
            const uint64_t hi = (lo + w[k]) < P ? (lo + w[k]) : P;
            const uint64_t csz = 1ull << chunk_shift;
            for (uint64_t c = (lo + csz - 1) >> chunk_shift; (c << chunk_shift) < hi; ++c)
                chunk_start[c] = r;
            ++r;
            off += w[k];
        }
    }
}

// ------------------------------------------------------------------ splat
__device__ __forceinline__ SplatRecord load_record(const SplatRecord* __restrict__ recs, uint32_t g) {
    const uint4* rp = reinterpret_cast<const uint4*>(recs + g);
    const uint4 a = rp[0], b = rp[1], cc = rp[2], d = rp[3];
    SplatRecord rec;
    rec.pmx = __uint_as_float(a.x);
    rec.pmy = __uint_as_float(a.y);
    rec.c0 = __uint_as_float(a.z);
    rec.c1 = __uint_as_float(a.w);
    rec.c2 = __uint_as_float(b.x);
    rec.conx = __uint_as_float(b.y);
    rec.cony = __uint_as_float(b.z);
    rec.conz = __uint_as_float(b.w);
    rec.opacity = __uint_as_float(cc.x);
    rec.inv_alpha = __uint_as_float(cc.y);
    rec.dilog_alpha = __uint_as_float(cc.z);
    rec.bmin = cc.w;
    rec.bmax = d.x;
    rec.packed_lo = d.y;
    rec.packed_hi = d.z;
    rec.alpha_t = __uint_as_float(d.w);
    return rec;
}

// The alpha test of a sample (gaussian_point_splatting.cu:118-123), decided
// from the exponent against the record's threshold and evaluated with the
// exponential only near it (alpha_test_banded: the same decision).
__device__ __forceinline__ bool alpha_passes(const SplatRecord& rec, float sigma) {
    return alpha_test_banded(rec.opacity, rec.alpha_t, sigma);
}

// One sample of the biased splat (use_unbiased_2d_splatting = 0) on its
// framebuffer word, as gaussian_point_splatting.cu:130-138 does it: the
// pre-check (a stored word that is nearer, or at the same depth with colour
// 0, skips the sample), then atomicAddColor (post_processing_kernels.cuh:13-42)
// as a 64-bit CAS loop: a nearer sample replaces the word, a farther one
// leaves it, one at equal depth adds its 36 colour bits modulo 2^36.  The
// pre-check's read is the loop's first guess.  Returns 1 if the loop ran.
__device__ __forceinline__ uint32_t add_color_sample(unsigned long long* __restrict__ w, unsigned long long value) {
    const unsigned long long depth = value & GPS_DEPTH_MASK;
    unsigned long long old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old <= depth) return 0u;
    unsigned long long assumed;
    do {
        assumed = old;
        const unsigned long long assumed_depth = assumed & GPS_DEPTH_MASK;
        if (depth > assumed_depth) break;
        const unsigned long long nv =
            depth == assumed_depth ? (assumed_depth | ((assumed + value) & GPS_COLOR_MASK)) : value;
        old = atomicCAS(w, assumed, nv);
    } while (old != assumed);
    return 1u;
}

// The K * spp samples of point p (seed stream p of RNG frame `frame_seed`)
// of one Gaussian (splatGaussianPoint, gaussian_point_splatting.cu:67-141):
// sample q = layer * K + k lands in layer q / K.  Returns the atomics issued.
__device__ __forceinline__ uint32_t splat_point(const SplatRecord& rec, uint32_t p, uint32_t frame_seed, int K,
                                                int spp, int up_w, int unbiased, unsigned long long* __restrict__ fb,
                                                int& last, uint32_t* s_key = nullptr,
                                                unsigned long long* s_val = nullptr, uint32_t key_base = 0u) {
    const int bminx = (int)(rec.bmin & 0xFFFFu), bminy = (int)(rec.bmin >> 16);
    const int bmaxx = (int)(rec.bmax & 0xFFFFu), bmaxy = (int)(rec.bmax >> 16);
    const unsigned long long packed = ((unsigned long long)rec.packed_hi << 32) | rec.packed_lo;
    const unsigned long long depth_bits = packed & ~(unsigned long long)GPS_COLOR_MASK;
    const int samples = spp * K;
    uint32_t issued = 0;
    u32x2 seed = make_seed(p, frame_seed);
    for (int layer = 0, k = 0, q = 0; q < samples; ++q) {
        if (k == K) {
            k = 0;
            ++layer;
        }
        ++k;
        const f32x2 u = pcg2d(seed);
        const f32x2 xy = unbiased ? corrected_box_muller_pre(u.x, u.y, rec.inv_alpha, rec.dilog_alpha)
                                  : box_muller(u.x, u.y);
        const float fx = det_floor(rec.pmx + rec.c0 * xy.x);
        const float fy = det_floor((rec.pmy + rec.c1 * xy.x) + rec.c2 * xy.y);
        int x, y;  // det_f2i_rz(fx), det_f2i_rz(fy) inside the bounds
        if (!det_pixel_in(fx, bminx, bmaxx, &x) || !det_pixel_in(fy, bminy, bmaxy, &y)) continue;
        const float dx = rec.pmx - fx, dy = rec.pmy - fy;
        const float sigma = -0.5f * (rec.conx * dx * dx + rec.conz * dy * dy) - rec.cony * dx * dy;
        if (!alpha_passes(rec, sigma)) continue;
        const int idx = fb_index(x, y, layer, spp, up_w);
        if (!unbiased) {
            // every sample adds: no repeat skip, no word table
            issued += add_color_sample(fb + idx, packed);
            continue;
        }
        // The word this thread just read or wrote for the same Gaussian
        // (same packed value): a second atomicMin would change nothing.
        if (idx == last) continue;
        last = idx;
#if GPS_SPLAT_SLOTS
        if (s_key) {
            // merge into the chunk's LDS word table when the slot is free or
            // already holds this word; otherwise go to the framebuffer
            // key: the word's index from the table's framebuffer base
            const uint32_t word = key_base + (uint32_t)idx;
            // Morton-adjacent words stay in nearby slots (a multiplicative
            // hash measured 1.195 ms against 0.913)
            const uint32_t slot = (word ^ (word >> 11)) & (GPS_SPLAT_SLOTS - 1u);
            uint32_t key = s_key[slot];
            if (key == 0xFFFFFFFFu) {
                const uint32_t old = atomicCAS(&s_key[slot], 0xFFFFFFFFu, word);
                key = old == 0xFFFFFFFFu ? word : old;
            }
            if (key == word) {
                atomicMin(&s_val[slot], packed);
                continue;
            }
#if GPS_SPLAT_PROBE2
            const uint32_t slot2 = slot ^ 1u;
            key = s_key[slot2];
            if (key == 0xFFFFFFFFu) {
                const uint32_t old = atomicCAS(&s_key[slot2], 0xFFFFFFFFu, word);
                key = old == 0xFFFFFFFFu ? word : old;
            }
            if (key == word) {
                atomicMin(&s_val[slot2], packed);
                continue;
            }
#if GPS_SPLAT_PROBE3
            const uint32_t slot3 = slot ^ 2u;
            key = s_key[slot3];
            if (key == 0xFFFFFFFFu) {
                const uint32_t old = atomicCAS(&s_key[slot3], 0xFFFFFFFFu, word);
                key = old == 0xFFFFFFFFu ? word : old;
            }
            if (key == word) {
                atomicMin(&s_val[slot3], packed);
                continue;
            }
#endif
#endif
        }
#endif
        if (fb[idx] <= depth_bits) continue;
        atomicMin(fb + idx, packed);
        ++issued;
    }
    return issued;
}


#if GPS_SPLAT_SLOTS
// End of a chunk: every stored word is read once and lowered with one
// atomicMin if the chunk's minimum is smaller, and the table is cleared.  A
// thread's reads are all issued before any is used (empty slots read word 0,
// which is always allocated), so the chunk waits for one memory round trip
// instead of one per slot.
template <int THREADS>
__device__ __forceinline__ uint32_t flush_word_table(uint32_t* s_key, unsigned long long* s_val,
                                                     unsigned long long* __restrict__ fb) {
    constexpr int PER = GPS_SPLAT_SLOTS / THREADS;
    static_assert(PER * THREADS == GPS_SPLAT_SLOTS, "slots split evenly over the threads");
    const int tid = threadIdx.x;
    uint32_t key[PER];
    unsigned long long cur[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) key[k] = s_key[tid + k * THREADS];
#pragma unroll
    for (int k = 0; k < PER; ++k) cur[k] = fb[key[k] != 0xFFFFFFFFu ? key[k] : 0u];
    uint32_t issued = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (key[k] == 0xFFFFFFFFu) continue;
        const unsigned long long v = s_val[tid + k * THREADS];
        if (cur[k] > v) {
            atomicMin(fb + key[k], v);
            ++issued;
        }
        s_key[tid + k * THREADS] = 0xFFFFFFFFu;
        s_val[tid + k * THREADS] = ~0ull;
    }
    return issued;
}
#endif

// A Gaussian whose bounds (16-pixel aligned) cover at most
// GPS_ORD_MAX_TILES 16x16 tiles whose maximum depth (tmax, the tiles in
// framebuffer order: 64x64 tiles of tw64 per row, 16 Morton-ordered 16x16
// tiles in each) is below its own: each of its samples lands inside the
// bounds on a word that is already nearer, so none can lower a word.
__device__ __forceinline__ bool tile_behind(const SplatRecord& rec, const uint32_t* __restrict__ tmax, int tw64) {
    const uint32_t gd = rec.packed_hi >> 4;  // depth28 of the packed word
    const int tx0 = (int)(rec.bmin & 0xFFFFu) >> 4, ty0 = (int)(rec.bmin >> 20);
    const int tx1 = (int)(rec.bmax & 0xFFFFu) >> 4, ty1 = (int)(rec.bmax >> 20);
    if ((tx1 - tx0) * (ty1 - ty0) > GPS_ORD_MAX_TILES) return false;
    for (int ty = ty0; ty < ty1; ++ty)
        for (int tx = tx0; tx < tx1; ++tx) {
            const uint32_t t = ((uint32_t)((ty >> 2) * tw64 + (tx >> 2)) << 4) |
                               (uint32_t)((tx & 1) | ((ty & 1) << 1) | ((tx & 2) << 1) | ((ty & 2) << 2));
            if (tmax[t] >= gd) return false;
        }
    return true;
}

// One chunk of the point-parallel splat: the chunk's entry offsets into LDS,
// PPT consecutive points per thread (a binary search for the first one),
// then the word table's flush.  With TEST and tmax set, the points of a
// Gaussian that tile_behind() clears are skipped and counted in `skipped`.
template <bool TEST>
__device__ __forceinline__ uint32_t splat_chunk(uint32_t c, const SplatRecord* __restrict__ recs,
                                                const uint32_t* __restrict__ cidx, const uint32_t* __restrict__ coff,
                                                const uint32_t* __restrict__ chunk_start, uint32_t P, uint32_t V,
                                                unsigned long long* __restrict__ fb, const SplatConsts& sc,
                                                uint32_t* s_off, uint32_t* s_key, unsigned long long* s_val,
                                                const uint32_t* __restrict__ tmax, int tw64, uint32_t& skipped) {
    const int tid = threadIdx.x;
    uint32_t issued = 0;
    const uint32_t r_lo = chunk_start[c];
    for (int j = tid; j <= GPS_SPLAT_CHUNK; j += GPS_SPLAT_THREADS) {
        const uint32_t r = r_lo + (uint32_t)j;
        s_off[j] = r < V ? coff[r] : 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint32_t p0 = c * GPS_SPLAT_CHUNK + (uint32_t)tid * GPS_SPLAT_PPT;
    if (p0 < P) {
        // last j with s_off[j] <= p0 (s_off[0] <= c*CHUNK <= p0)
        int lo = 0, hi = GPS_SPLAT_CHUNK;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= p0) lo = mid;
            else hi = mid - 1;
        }
        int j = lo;
        uint32_t cur_g = 0xFFFFFFFFu;
        SplatRecord rec;
        int last = -1;
        bool skip = false;
        for (int k = 0; k < GPS_SPLAT_PPT; ++k) {
            const uint32_t p = p0 + (uint32_t)k;
            if (p >= P) break;
            while (j < GPS_SPLAT_CHUNK - 1 && s_off[j + 1] <= p) ++j;
            if (r_lo + (uint32_t)j >= V) break;
            const uint32_t g = cidx[r_lo + (uint32_t)j];
            if (g != cur_g) {
                cur_g = g;
                last = -1;
                if (TEST && tmax) {
                    // the test's words first (bounds, depth: the record's
                    // second half); the rest only for a Gaussian that stays
                    const uint4* rp = reinterpret_cast<const uint4*>(recs + g);
                    const uint4 cc = rp[2], d = rp[3];
                    rec.bmin = cc.w;
                    rec.bmax = d.x;
                    rec.packed_hi = d.z;
                    skip = tile_behind(rec, tmax, tw64);
                    if (!skip) rec = load_record(recs, g);
                } else {
                    rec = load_record(recs, g);
                }
            }
            if (TEST && skip) {
                ++skipped;
                continue;
            }
            issued += splat_point(rec, p, sc.frame_seed, sc.K, sc.spp, sc.up_w, sc.unbiased, fb, last, s_key, s_val);
        }
    }
    __syncthreads();
#if GPS_SPLAT_SLOTS
    // flush the chunk's merged words and clear the table; the next chunk
    // inserts only after the barrier that follows its offsets
    issued += flush_word_table<GPS_SPLAT_THREADS>(s_key, s_val, fb);
#endif
    return issued;
}

__global__ __launch_bounds__(GPS_SPLAT_THREADS) void splat_kernel(const SplatRecord* __restrict__ recs,
                                                                  const uint32_t* __restrict__ cidx,
                                                                  const uint32_t* __restrict__ coff,
                                                                  const uint32_t* __restrict__ chunk_start,
                                                                  const PassCounters* __restrict__ ctr,
                                                                  unsigned long long* __restrict__ fb, SplatConsts sc,
                                                                  uint32_t* __restrict__ atomics_count,
                                                                  const uint32_t* __restrict__ only_if) {
    __shared__ uint32_t s_off[GPS_SPLAT_CHUNK + 1];
#if GPS_SPLAT_SLOTS
    __shared__ uint32_t s_key[GPS_SPLAT_SLOTS];
    __shared__ unsigned long long s_val[GPS_SPLAT_SLOTS];
#else
    uint32_t* const s_key = nullptr;
    unsigned long long* const s_val = nullptr;
#endif
    if (only_if && *only_if == 0u) return;  // the binned splat covers this pass
    const uint32_t P = ctr->points;
    const uint32_t V = ctr->valid;
    const uint32_t nchunks = ctr->chunks;
    const int tid = threadIdx.x;
    uint32_t issued = 0, skipped = 0;
#if GPS_SPLAT_SLOTS
    for (int i = tid; i < GPS_SPLAT_SLOTS; i += GPS_SPLAT_THREADS) {
        s_key[i] = 0xFFFFFFFFu;
        s_val[i] = ~0ull;
    }  // ordered before the first insert by the barrier after the offsets
#endif
    // Workgroups are dispatched round-robin over the XCDs (blockIdx % xcds);
    // give every XCD one contiguous range of chunks so the framebuffer lines
    // of a compact scene region stay in that XCD's L2.
    const uint32_t xcds = (uint32_t)sc.xcds;
    const uint32_t xcd = blockIdx.x % xcds;
    const uint32_t per_xcd = gridDim.x / xcds;
    const uint32_t seg_lo = (uint32_t)(((uint64_t)nchunks * xcd) / xcds);
    const uint32_t seg_hi = (uint32_t)(((uint64_t)nchunks * (xcd + 1)) / xcds);
    for (uint32_t c = seg_lo + blockIdx.x / xcds; c < seg_hi; c += per_xcd)
        issued += splat_chunk<false>(c, recs, cidx, coff, chunk_start, P, V, fb, sc, s_off, s_key, s_val, nullptr, 0,
                                     skipped);
    if (atomics_count) add_count(atomics_count, issued);
}

// ------------------------------------------------------------------ ordered splat
// The ordered form of the point-parallel splat for dense passes (see
// GPS_ORD_BINS): ord_key_kernel buckets the chunks by depth, ord_scan_kernel
// turns the bucket histogram into the phases' bucket limits, and every phase
// splats its chunks after ord_tile_max_kernel has taken the tiles' maximum
// depth.  Every kernel first evaluates ord_active on the pass's counters (the
// same decision everywhere); an inactive pass is splatted by phase 0 in the
// plain chunk order, the other launches return at once.
struct OrdGate {
    uint32_t K, up_w, up_h;
    int32_t forced;  // order whatever the density (tests)
};

__device__ __forceinline__ bool ord_active(const PassCounters* __restrict__ ctr, const OrdGate& g) {
    const uint32_t n = ctr->chunks;
    if (n == 0u) return false;
    if (g.forced) return true;
    return n >= GPS_ORD_MIN_CHUNKS &&
           (uint64_t)ctr->points * g.K > (uint64_t)GPS_ORD_DENSE_SAMPLES * g.up_w * (uint64_t)g.up_h;
}

// Log-depth bucket: depths below 128 one bucket each, then 128 buckets per
// power of two (monotone; 2816 buckets for 28-bit depths).
__device__ __forceinline__ uint32_t ord_bucket(uint32_t d) {
    if (d < 128u) return d;
    const uint32_t e = 31u - (uint32_t)__builtin_clz(d);  // 7..27
    return (e - 6u) * 128u + ((d >> (e - 7u)) & 127u);
}

__global__ __launch_bounds__(256) void ord_key_kernel(const SplatRecord* __restrict__ recs,
                                                      const uint32_t* __restrict__ cidx,
                                                      const uint32_t* __restrict__ chunk_start,
                                                      const PassCounters* __restrict__ ctr, OrdGate gate,
                                                      const uint32_t* __restrict__ only_if,
                                                      uint16_t* __restrict__ bucket, uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_h[GPS_ORD_BINS];
    if (only_if && *only_if == 0u) return;
    if (!ord_active(ctr, gate)) return;
    const uint32_t n = ctr->chunks;
    for (uint32_t i = threadIdx.x; i < GPS_ORD_BINS; i += 256) s_h[i] = 0u;
    __syncthreads();
    for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < n; c += gridDim.x * 256u) {
        const uint32_t g = cidx[chunk_start[c]];
        const uint32_t d = reinterpret_cast<const uint32_t*>(recs + g)[14] >> 4;  // packed_hi: depth28
        const uint32_t b = ord_bucket(d);
        bucket[c] = (uint16_t)b;
        atomicAdd(&s_h[b], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < GPS_ORD_BINS; i += 256)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// One workgroup: lim[k] = one past the bucket in which the running count of
// chunks reaches n / 2^(GPS_ORD_PHASES - 1 - k) (n/8, n/4, n/2, n for four
// phases); clears the histogram.
__global__ __launch_bounds__(1024) void ord_scan_kernel(const PassCounters* __restrict__ ctr, OrdGate gate,
                                                       const uint32_t* __restrict__ only_if,
                                                       uint32_t* __restrict__ hist, uint32_t* __restrict__ lim) {
    static_assert(GPS_ORD_PHASES >= 2 && GPS_ORD_PHASES <= 8 && 3u * 1024u >= GPS_ORD_BINS,
                  "three buckets per thread, phases of n / 2^k");
    __shared__ uint32_t s_w[16];
    if (only_if && *only_if == 0u) return;
    if (!ord_active(ctr, gate)) return;
    const uint32_t n = ctr->chunks;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t h[3], sum = 0u;
    for (int k = 0; k < 3; ++k) {
        const uint32_t b = (uint32_t)tid * 3u + (uint32_t)k;
        h[k] = b < GPS_ORD_BINS ? hist[b] : 0u;
        sum += h[k];
    }
    const uint32_t incl = wave_scan_incl(sum);
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (int w = 0; w < wave; ++w) run += s_w[w];
    uint32_t t[GPS_ORD_PHASES];
    for (int ph = 0; ph < GPS_ORD_PHASES - 1; ++ph) t[ph] = max(1u, n >> (GPS_ORD_PHASES - 1 - ph));
    t[GPS_ORD_PHASES - 1] = n;
    for (int k = 0; k < 3; ++k) {
        const uint32_t b = (uint32_t)tid * 3u + (uint32_t)k;
        if (b >= GPS_ORD_BINS) break;
        const uint32_t excl = run;
        run += h[k];
        for (int ph = 0; ph < GPS_ORD_PHASES; ++ph)
            if (excl < t[ph] && run >= t[ph]) lim[ph] = b + 1u;
        hist[b] = 0u;
    }
}

// Maximum depth28 of every 16x16-pixel tile (256 * spp consecutive words in
// the framebuffer's Morton layout), one wave per tile.
__global__ __launch_bounds__(256) void ord_tile_max_kernel(const unsigned long long* __restrict__ fb,
                                                           uint32_t* __restrict__ tmax, uint32_t ntiles, int spp,
                                                           const PassCounters* __restrict__ ctr, OrdGate gate,
                                                           const uint32_t* __restrict__ only_if) {
    if (only_if && *only_if == 0u) return;
    if (!ord_active(ctr, gate)) return;
    const uint32_t t = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= ntiles) return;
    const uint4* w = reinterpret_cast<const uint4*>(fb + (uint64_t)t * 256u * (uint32_t)spp);
    uint32_t m = 0u;
    for (int i = lane; i < 128 * spp; i += 64) {
        const uint4 v = w[i];  // two words: (lo, hi) pairs
        m = max(m, max(v.y >> 4, v.w >> 4));
    }
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
    if (lane == 0) tmax[t] = m;
}

// One phase of the ordered splat: the chunks whose bucket is in
// [lim[phase-1], lim[phase]), in groups of GPS_ORD_GROUP consecutive chunks
// dealt round-robin to the XCDs; with `test`, Gaussians behind the tiles are
// skipped.  An inactive pass is splatted by phase 0 in splat_kernel's order.
__global__ __launch_bounds__(GPS_SPLAT_THREADS) void splat_ordered_kernel(
    const SplatRecord* __restrict__ recs, const uint32_t* __restrict__ cidx, const uint32_t* __restrict__ coff,
    const uint32_t* __restrict__ chunk_start, const PassCounters* __restrict__ ctr, unsigned long long* __restrict__ fb,
    SplatConsts sc, uint32_t* __restrict__ atomics_count, const uint32_t* __restrict__ only_if,
    const uint16_t* __restrict__ bucket, const uint32_t* __restrict__ lim, const uint32_t* __restrict__ tmax,
    int tw64, int phase, int test, OrdGate gate, uint32_t* __restrict__ skipped_count) {
    __shared__ uint32_t s_off[GPS_SPLAT_CHUNK + 1];
#if GPS_SPLAT_SLOTS
    __shared__ uint32_t s_key[GPS_SPLAT_SLOTS];
    __shared__ unsigned long long s_val[GPS_SPLAT_SLOTS];
#else
    uint32_t* const s_key = nullptr;
    unsigned long long* const s_val = nullptr;
#endif
    if (only_if && *only_if == 0u) return;
    const bool active = ord_active(ctr, gate);
    if (!active && phase > 0) return;
    const uint32_t P = ctr->points;
    const uint32_t V = ctr->valid;
    const uint32_t nchunks = ctr->chunks;
    const int tid = threadIdx.x;
    uint32_t issued = 0, skipped = 0;
#if GPS_SPLAT_SLOTS
    for (int i = tid; i < GPS_SPLAT_SLOTS; i += GPS_SPLAT_THREADS) {
        s_key[i] = 0xFFFFFFFFu;
        s_val[i] = ~0ull;
    }
#endif
    const uint32_t xcds = (uint32_t)sc.xcds;
    const uint32_t xcd = blockIdx.x % xcds;
    const uint32_t per_xcd = gridDim.x / xcds;
    if (!active) {
        const uint32_t seg_lo = (uint32_t)(((uint64_t)nchunks * xcd) / xcds);
        const uint32_t seg_hi = (uint32_t)(((uint64_t)nchunks * (xcd + 1)) / xcds);
        for (uint32_t c = seg_lo + blockIdx.x / xcds; c < seg_hi; c += per_xcd)
            issued += splat_chunk<true>(c, recs, cidx, coff, chunk_start, P, V, fb, sc, s_off, s_key, s_val, nullptr,
                                        0, skipped);
    } else {
        const uint32_t lo = phase > 0 ? lim[phase - 1] : 0u, hi = lim[phase];
        const uint32_t ngroups = (nchunks + GPS_ORD_GROUP - 1u) / GPS_ORD_GROUP;
        const uint32_t mine = xcd < ngroups ? (ngroups - xcd + xcds - 1u) / xcds : 0u;  // groups of this XCD
        const uint32_t items = mine * GPS_ORD_GROUP;
        const uint32_t* const tm = test ? tmax : nullptr;
        for (uint32_t i = blockIdx.x / xcds; i < items; i += per_xcd) {
            const uint32_t c = ((i / GPS_ORD_GROUP) * xcds + xcd) * GPS_ORD_GROUP + i % GPS_ORD_GROUP;
            if (c >= nchunks) continue;
            const uint32_t b = bucket[c];
            if (b < lo || b >= hi) continue;
            issued += splat_chunk<true>(c, recs, cidx, coff, chunk_start, P, V, fb, sc, s_off, s_key, s_val, tm, tw64,
                                        skipped);
        }
    }
    if (atomics_count) add_count(atomics_count, issued);
    if (skipped_count) add_count(skipped_count, skipped);
}

// ------------------------------------------------------------------ binned splat
// Screen regions of at most GPS_BIN_WORDS framebuffer words (128x128 px at
// 1 spp, 128x64 at 2, 64x64 at 3-4) are rendered by one workgroup in LDS.
// The valid Gaussians are binned to the regions their bounds overlap; a
// region job replays the points of its Gaussians with their global point
// index (hence exactly the samples of the point-parallel splat) and keeps
// the samples that land inside the region, with LDS atomicMin.  A region
// rendered by one job starts from the framebuffer and stores the region
// back; a heavy region split into several jobs merges its changed words with
// global atomicMin.  Gaussians overlapping more than GPS_BIN_MAX_REGIONS
// regions are splatted afterwards with global atomics (no replay).  Since
// min is order-independent the framebuffer is bit-identical to the
// point-parallel splat.  If the region lists overflow their allocation
// (state[3] = 1), or the pass has more points than dense_points (state[3] = 2,
// the automatic choice's density test), the binned kernels stand down and the
// point-parallel splat runs instead.

__device__ __forceinline__ uint32_t bin_region_words(const BinGeom& g) {
    return (uint32_t)(g.rtx * g.rty) * GPS_TILE_WORDS * (uint32_t)g.spp;
}

// Entries [r0, r1) of one frame whose entries end at v_end, whose points end
// at t_end and whose point budget ends at p_end: points per entry (clipped to
// the budget), region rectangle, region histogram in LDS (a workgroup takes
// one contiguous entry range, so the Morton-ordered entries touch few
// regions), big entries to their list.  The histogram is added to the
// frame's regions, which start at rbase.
__device__ __forceinline__ void bin_count_range(const SplatRecord* __restrict__ recs, const uint32_t* __restrict__ cidx,
                                                const uint32_t* __restrict__ coff, const BinGeom& g, uint32_t r0,
                                                uint32_t r1, uint32_t v_end, uint32_t t_end, uint32_t p_end,
                                                uint32_t rbase, uint2* __restrict__ erect, uint32_t* __restrict__ ew,
                                                uint32_t* __restrict__ reg_cnt, uint32_t* __restrict__ reg_pts,
                                                uint32_t* __restrict__ big, uint32_t* __restrict__ state,
                                                uint32_t* s_cnt, uint32_t* s_pts) {
    const int tid = threadIdx.x;
    for (uint32_t i = tid; i < g.nregions; i += 256) {
        s_cnt[i] = 0u;
        s_pts[i] = 0u;
    }
    __syncthreads();
    for (uint32_t r = r0 + tid; r < r1; r += 256) {
        const uint32_t lo = coff[r];
        const uint32_t hi = min(r + 1 < v_end ? coff[r + 1] : t_end, p_end);
        uint32_t w = lo < hi ? hi - lo : 0u;
        uint2 rect = make_uint2(0xFFFFFFFFu, 0u);
        if (w) {
            const uint32_t* rp = reinterpret_cast<const uint32_t*>(recs + cidx[r]);
            const uint32_t bmin = rp[11], bmax = rp[12];
            const uint32_t bx0 = bmin & 0xFFFFu, by0 = bmin >> 16, bx1 = bmax & 0xFFFFu, by1 = bmax >> 16;
            if (bx1 > bx0 && by1 > by0) {
                const uint32_t rx0 = bx0 >> g.rw_shift, rx1 = (bx1 - 1u) >> g.rw_shift;
                const uint32_t ry0 = by0 >> g.rh_shift, ry1 = (by1 - 1u) >> g.rh_shift;
                if ((rx1 - rx0 + 1u) * (ry1 - ry0 + 1u) > GPS_BIN_MAX_REGIONS) {
                    big[atomicAdd(&state[2], 1u)] = r;
                } else {
                    rect = make_uint2(rx0 | (rx1 << 16), ry0 | (ry1 << 16));
                    for (uint32_t ry = ry0; ry <= ry1; ++ry)
                        for (uint32_t rx = rx0; rx <= rx1; ++rx) {
                            const uint32_t reg = ry * (uint32_t)g.nrx + rx;
                            atomicAdd(&s_cnt[reg], 1u);
                            atomicAdd(&s_pts[reg], w);
                        }
                }
            } else {
                w = 0u;  // empty bounds: every sample would be rejected
            }
        }
        erect[r] = rect;
        ew[r] = w;
    }
    __syncthreads();
    for (uint32_t i = tid; i < g.nregions; i += 256)
        if (s_cnt[i]) {
            atomicAdd(&reg_cnt[rbase + i], s_cnt[i]);
            atomicAdd(&reg_pts[rbase + i], s_pts[i]);
        }
}

__global__ __launch_bounds__(256) void bin_count_kernel(const SplatRecord* __restrict__ recs,
                                                        const uint32_t* __restrict__ cidx,
                                                        const uint32_t* __restrict__ coff,
                                                        const PassCounters* __restrict__ ctr, BinGeom g,
                                                        uint2* __restrict__ erect, uint32_t* __restrict__ ew,
                                                        uint32_t* __restrict__ reg_cnt, uint32_t* __restrict__ reg_pts,
                                                        uint32_t* __restrict__ big, uint32_t* __restrict__ state,
                                                        uint32_t dense_points) {
    extern __shared__ uint32_t s_dyn[];  // 2 * nregions words
    const uint32_t V = ctr->valid, P = ctr->points, T = ctr->total;
    if (P > dense_points) return;  // too dense to bin: bin_scan_kernel stands the binned kernels down
    const uint32_t per = (V + gridDim.x - 1) / gridDim.x;
    const uint32_t r0 = min(V, blockIdx.x * per), r1 = min(V, r0 + per);
    bin_count_range(recs, cidx, coff, g, r0, r1, V, T, P, 0u, erect, ew, reg_cnt, reg_pts, big, state, s_dyn,
                    s_dyn + g.nregions);
}
// One workgroup: region offsets, jobs of at most `job_points` points per
// region, the job table, and the list length (overflow flag).  Clears the
// region counters for the next pass.
__global__ __launch_bounds__(1024) void bin_scan_kernel(BinGeom g, uint32_t* __restrict__ reg_cnt,
                                                        uint32_t* __restrict__ reg_pts,
                                                        uint32_t* __restrict__ reg_ecnt,
                                                        uint32_t* __restrict__ reg_ptot,
                                                        uint32_t* __restrict__ reg_eoff,
                                                        uint32_t* __restrict__ reg_poff,
                                                        uint32_t* __restrict__ reg_cursor,
                                                        uint32_t* __restrict__ reg_jobs,
                                                        uint32_t* __restrict__ job_reg, uint32_t* __restrict__ job_p0,
                                                        uint32_t* __restrict__ state, uint32_t list_cap,
                                                        uint32_t job_points, uint32_t job_cap,
                                                        const PassCounters* __restrict__ ctr, uint32_t dense_points) {
    __shared__ uint32_t s_a[16], s_b[16], s_c[16];
    __shared__ uint32_t s_ca, s_cb, s_cc;
    __shared__ unsigned long long s_wide_e, s_wide_p;  // 64-bit totals: list and point overflow
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        s_ca = 0u;
        s_cb = 0u;
        s_cc = 0u;
        s_wide_e = 0ull;
        s_wide_p = 0ull;
    }
    __syncthreads();
    for (uint32_t base = 0; base < g.nregions; base += 1024) {
        const uint32_t i = base + tid;
        const uint32_t cnt = i < g.nregions ? reg_cnt[i] : 0u;
        const uint32_t pts = i < g.nregions ? reg_pts[i] : 0u;
        if (cnt) {
            atomicAdd(&s_wide_e, (unsigned long long)cnt);
            atomicAdd(&s_wide_p, (unsigned long long)pts);
        }
        const uint32_t nj = pts ? (pts + job_points - 1u) / job_points : 0u;
        const uint32_t ia = wave_scan_incl(cnt), ib = wave_scan_incl(pts), ic = wave_scan_incl(nj);
        if (lane == 63) {
            s_a[wave] = ia;
            s_b[wave] = ib;
            s_c[wave] = ic;
        }
        __syncthreads();
        uint32_t wa = 0, wb = 0, wc = 0;
        for (int w = 0; w < wave; ++w) {
            wa += s_a[w];
            wb += s_b[w];
            wc += s_c[w];
        }
        const uint32_t ca = s_ca, cb = s_cb, cc = s_cc;
        if (i < g.nregions) {
            const uint32_t eoff = ca + wa + ia - cnt, joff = cc + wc + ic - nj;
            reg_ecnt[i] = cnt;
            reg_ptot[i] = pts;
            reg_eoff[i] = eoff;
            reg_poff[i] = cb + wb + ib - pts;
            reg_cursor[i] = eoff;
            reg_jobs[i] = nj;
            for (uint32_t j = 0; j < nj && joff + j < job_cap; ++j) {
                job_reg[joff + j] = i;
                job_p0[joff + j] = j * job_points;
            }
            reg_cnt[i] = 0u;
            reg_pts[i] = 0u;
        }
        __syncthreads();
        if (tid == 1023) {
            s_ca = ca + wa + ia;
            s_cb = cb + wb + ib;
            s_cc = cc + wc + ic;
        }
        __syncthreads();
    }
    if (tid == 0) {
        state[0] = s_ca;
        state[1] = min(s_cc, job_cap);
        const bool dense = ctr && ctr->points > dense_points;
        state[3] = dense ? 2u : (s_wide_e > list_cap || s_cc > job_cap || s_wide_p > 0xFFFFFFFFull) ? 1u : 0u;
        state[4] = 0u;
        state[5] = 0u;
    }
}

// Region lists: the entry ranges of bin_count_range; the workgroup counts its
// entries per region in LDS, reserves one range per region with a single
// global atomic, and fills it.
__device__ __forceinline__ void bin_scatter_range(const BinGeom& g, uint32_t r0, uint32_t r1, uint32_t rbase,
                                                  const uint2* __restrict__ erect, const uint32_t* __restrict__ ew,
                                                  uint32_t* __restrict__ reg_cursor, uint32_t* __restrict__ list,
                                                  uint32_t* __restrict__ lw, uint32_t* s_cnt, uint32_t* s_base) {
    const int tid = threadIdx.x;
    for (uint32_t i = tid; i < g.nregions; i += 256) s_cnt[i] = 0u;
    __syncthreads();
    for (uint32_t r = r0 + tid; r < r1; r += 256) {
        const uint2 rect = erect[r];
        if (rect.x == 0xFFFFFFFFu) continue;
        for (uint32_t ry = rect.y & 0xFFFFu; ry <= (rect.y >> 16); ++ry)
            for (uint32_t rx = rect.x & 0xFFFFu; rx <= (rect.x >> 16); ++rx)
                atomicAdd(&s_cnt[ry * (uint32_t)g.nrx + rx], 1u);
    }
    __syncthreads();
    for (uint32_t i = tid; i < g.nregions; i += 256) {
        if (s_cnt[i]) s_base[i] = atomicAdd(&reg_cursor[rbase + i], s_cnt[i]);
        s_cnt[i] = 0u;
    }
    __syncthreads();
    for (uint32_t r = r0 + tid; r < r1; r += 256) {
        const uint2 rect = erect[r];
        if (rect.x == 0xFFFFFFFFu) continue;
        const uint32_t w = ew[r];
        for (uint32_t ry = rect.y & 0xFFFFu; ry <= (rect.y >> 16); ++ry)
            for (uint32_t rx = rect.x & 0xFFFFu; rx <= (rect.x >> 16); ++rx) {
                const uint32_t reg = ry * (uint32_t)g.nrx + rx;
                const uint32_t slot = s_base[reg] + atomicAdd(&s_cnt[reg], 1u);
                list[slot] = r;
                lw[slot] = w;
            }
    }
}

__global__ __launch_bounds__(256) void bin_scatter_kernel(const PassCounters* __restrict__ ctr, BinGeom g,
                                                          const uint2* __restrict__ erect,
                                                          const uint32_t* __restrict__ ew,
                                                          uint32_t* __restrict__ reg_cursor,
                                                          uint32_t* __restrict__ list, uint32_t* __restrict__ lw,
                                                          const uint32_t* __restrict__ state) {
    extern __shared__ uint32_t s_dyn[];  // 2 * nregions words
    if (state[3]) return;
    const uint32_t V = ctr->valid;
    const uint32_t per = (V + gridDim.x - 1) / gridDim.x;
    const uint32_t r0 = min(V, blockIdx.x * per), r1 = min(V, r0 + per);
    bin_scatter_range(g, r0, r1, 0u, erect, ew, reg_cursor, list, lw, s_dyn, s_dyn + g.nregions);
}
// Exclusive scan of lw over the list (region-major) in tiles of 4096 slots:
// per-tile sums, a one-workgroup scan of the tile sums, then the local scan.
// The list length comes from the device (grid-stride over its tiles).
__global__ __launch_bounds__(256) void bin_tile_sum_kernel(const uint32_t* __restrict__ lw,
                                                           const uint32_t* __restrict__ state,
                                                           uint32_t* __restrict__ tsum) {
    __shared__ uint32_t s[4];
    const uint32_t L = state[3] ? 0u : state[0];
    const uint32_t nt = (L + GPS_TILE - 1) / GPS_TILE;
    for (uint32_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
        const uint32_t t0 = tile * GPS_TILE + threadIdx.x * 16u;
        uint32_t sum = 0;
        if (t0 < L) {
            const uint4* p = reinterpret_cast<const uint4*>(lw + t0);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint4 v = p[k];
                const uint32_t b = t0 + 4u * k;
                sum += (b < L ? v.x : 0u) + (b + 1 < L ? v.y : 0u) + (b + 2 < L ? v.z : 0u) + (b + 3 < L ? v.w : 0u);
            }
        }
        sum = wave_reduce_add(sum);
        if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = sum;
        __syncthreads();
        if (threadIdx.x == 0) tsum[tile] = (s[0] + s[1]) + (s[2] + s[3]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void bin_tiles_scan_kernel(const uint32_t* __restrict__ tsum,
                                                              const uint32_t* __restrict__ state,
                                                              uint32_t* __restrict__ toff) {
    __shared__ uint32_t s_a[16];
    __shared__ uint32_t s_carry;
    const uint32_t L = state[3] ? 0u : state[0];
    const uint32_t nt = (L + GPS_TILE - 1) / GPS_TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_carry = 0u;
    __syncthreads();
    for (uint32_t base = 0; base < nt; base += 1024) {
        const uint32_t i = base + tid;
        const uint32_t a = i < nt ? tsum[i] : 0u;
        const uint32_t ia = wave_scan_incl(a);
        if (lane == 63) s_a[wave] = ia;
        __syncthreads();
        uint32_t wa = 0;
        for (int w = 0; w < wave; ++w) wa += s_a[w];
        const uint32_t c = s_carry;
        if (i < nt) toff[i] = c + wa + ia - a;
        __syncthreads();
        if (tid == 1023) s_carry = c + wa + ia;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void bin_local_scan_kernel(const uint32_t* __restrict__ lw,
                                                             const uint32_t* __restrict__ state,
                                                             const uint32_t* __restrict__ toff,
                                                             uint32_t* __restrict__ lpre) {
    __shared__ uint32_t s_a[4];
    const uint32_t L = state[3] ? 0u : state[0];
    const uint32_t nt = (L + GPS_TILE - 1) / GPS_TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
        const uint32_t b0 = tile * GPS_TILE + (uint32_t)tid * 16u;
        uint32_t w[16];
        const uint4* p = reinterpret_cast<const uint4*>(lw + b0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 v = b0 < L ? p[k] : make_uint4(0u, 0u, 0u, 0u);
            w[4 * k] = b0 + 4u * k < L ? v.x : 0u;
            w[4 * k + 1] = b0 + 4u * k + 1u < L ? v.y : 0u;
            w[4 * k + 2] = b0 + 4u * k + 2u < L ? v.z : 0u;
            w[4 * k + 3] = b0 + 4u * k + 3u < L ? v.w : 0u;
        }
        uint32_t sa = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) sa += w[k];
        const uint32_t ia = wave_scan_incl(sa);
        if (lane == 63) s_a[wave] = ia;
        __syncthreads();
        uint32_t wa = 0;
        for (int q = 0; q < wave; ++q) wa += s_a[q];
        uint32_t off = toff[tile] + wa + ia - sa;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (b0 + (uint32_t)k < L) lpre[b0 + k] = off;
            off += w[k];
        }
        __syncthreads();
    }
}

struct BinSplat {
    BinGeom g;
    int32_t K, unbiased;
    uint32_t frame_seed;
    uint64_t background;
};

// Samples of point p of one Gaussian that land in the LDS region.
__device__ __forceinline__ void splat_point_region(const SplatRecord& rec, uint32_t p, const BinSplat& b, int rx0,
                                                   int ry0, unsigned long long* s_fb) {
    const int bminx = (int)(rec.bmin & 0xFFFFu), bminy = (int)(rec.bmin >> 16);
    const int bmaxx = (int)(rec.bmax & 0xFFFFu), bmaxy = (int)(rec.bmax >> 16);
    const unsigned long long packed = ((unsigned long long)rec.packed_hi << 32) | rec.packed_lo;
    const unsigned long long depth_bits = packed & ~(unsigned long long)GPS_COLOR_MASK;
    const int rw = 1 << b.g.rw_shift, rh = 1 << b.g.rh_shift;
    const int K = b.K, spp = b.g.spp, samples = spp * K;
    u32x2 seed = make_seed(p, b.frame_seed);
    for (int layer = 0, k = 0, q = 0; q < samples; ++q) {
        if (k == K) {
            k = 0;
            ++layer;
        }
        ++k;
        const f32x2 u = pcg2d(seed);
        const f32x2 xy = b.unbiased ? corrected_box_muller_pre(u.x, u.y, rec.inv_alpha, rec.dilog_alpha)
                                    : box_muller(u.x, u.y);
        const float fx = det_floor(rec.pmx + rec.c0 * xy.x);
        const float fy = det_floor((rec.pmy + rec.c1 * xy.x) + rec.c2 * xy.y);
        int x, y;  // det_f2i_rz(fx), det_f2i_rz(fy) inside the bounds
        if (!det_pixel_in(fx, bminx, bmaxx, &x) || !det_pixel_in(fy, bminy, bmaxy, &y)) continue;
        const int lx = x - rx0, ly = y - ry0;
        if ((unsigned)lx >= (unsigned)rw || (unsigned)ly >= (unsigned)rh) continue;  // another region's sample
        const float dx = rec.pmx - fx, dy = rec.pmy - fy;
        const float sigma = -0.5f * (rec.conx * dx * dx + rec.conz * dy * dy) - rec.cony * dx * dy;
        if (!alpha_passes(rec, sigma)) continue;
        const uint32_t t = (uint32_t)((ly >> GPS_TILE_BITS) * b.g.rtx + (lx >> GPS_TILE_BITS));
        const uint32_t i = (t * GPS_TILE_WORDS + morton_encode((uint32_t)(x & (GPS_TILE_SIZE - 1)),
                                                                (uint32_t)(y & (GPS_TILE_SIZE - 1)))) *
                               (uint32_t)spp +
                           (uint32_t)layer;
        if (s_fb[i] <= depth_bits) continue;
        atomicMin(s_fb + i, packed);
    }
}

// One region job: the region's words from the framebuffer (a region with one
// job) or the background (several jobs), the job's points [p_lo, p_hi) of
// the region's list (e0, ecnt: its slots; pbase: its first point in list
// order), then the region back to the framebuffer: a plain store, or
// atomicMin of the changed words when the region has several jobs.  `pbias`
// turns the entries' point offsets into the frame's own point indices.
// Returns the global atomics issued.  Called by every thread of the
// workgroup.
__device__ __forceinline__ uint32_t bin_splat_job(const SplatRecord* __restrict__ recs,
                                                  const uint32_t* __restrict__ cidx, const uint32_t* __restrict__ coff,
                                                  const uint32_t* __restrict__ list, const uint32_t* __restrict__ lpre,
                                                  uint32_t e0, uint32_t ecnt, uint32_t pbase, uint32_t p_lo,
                                                  uint32_t p_hi, bool single, uint32_t lreg, const BinSplat& b,
                                                  unsigned long long* __restrict__ fb, uint32_t pbias,
                                                  unsigned long long* s_fb, uint32_t* s_off, uint32_t* s_slot0) {
    const int tid = threadIdx.x;
    const uint32_t words = bin_region_words(b.g);
    const uint32_t tile_words = GPS_TILE_WORDS * (uint32_t)b.g.spp;
    const int tiles_rows = (b.g.up_h + GPS_TILE_SIZE - 1) >> GPS_TILE_BITS;
    const int rxi = (int)(lreg % (uint32_t)b.g.nrx), ryi = (int)(lreg / (uint32_t)b.g.nrx);
    const int rx0 = rxi << b.g.rw_shift, ry0 = ryi << b.g.rh_shift;
    const int tx0 = rx0 >> GPS_TILE_BITS, ty0 = ry0 >> GPS_TILE_BITS;
    uint32_t merged = 0;
    // region words <- framebuffer (single job) or background (merged later)
    for (uint32_t w = (uint32_t)tid * 2u; w < words; w += 2u * GPS_BIN_THREADS) {
        const uint32_t t = w / tile_words, within = w - t * tile_words;
        const int tx = tx0 + (int)(t % (uint32_t)b.g.rtx), ty = ty0 + (int)(t / (uint32_t)b.g.rtx);
        uint4 v = make_uint4((uint32_t)b.background, (uint32_t)(b.background >> 32), (uint32_t)b.background,
                             (uint32_t)(b.background >> 32));
        if (single && tx < b.g.tiles_per_row && ty < tiles_rows)
            v = *reinterpret_cast<const uint4*>(fb + (uint64_t)(ty * b.g.tiles_per_row + tx) * tile_words + within);
        s_fb[w] = ((unsigned long long)v.y << 32) | v.x;
        s_fb[w + 1] = ((unsigned long long)v.w << 32) | v.z;
    }
    for (uint32_t q0 = p_lo; q0 < p_hi; q0 += GPS_BIN_CHUNK) {
        // last slot whose region-local prefix is <= q0
        if (tid == 0) {
            uint32_t lo = 0, hi = ecnt - 1u;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1u) >> 1;
                if (lpre[e0 + mid] - pbase <= q0) lo = mid;
                else hi = mid - 1u;
            }
            *s_slot0 = lo;
        }
        __syncthreads();
        const uint32_t j0 = *s_slot0;
        for (int j = tid; j <= GPS_BIN_CHUNK; j += GPS_BIN_THREADS)
            s_off[j] = j0 + (uint32_t)j < ecnt ? lpre[e0 + j0 + (uint32_t)j] - pbase : 0xFFFFFFFFu;
        __syncthreads();
        const uint32_t qa = q0 + (uint32_t)tid * GPS_BIN_PPT;
        if (qa < p_hi) {
            int lo = 0, hi = GPS_BIN_CHUNK;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= qa) lo = mid;
                else hi = mid - 1;
            }
            int j = lo;
            uint32_t cur = 0xFFFFFFFFu, pstart = 0;
            SplatRecord rec;
            for (int k = 0; k < GPS_BIN_PPT; ++k) {
                const uint32_t q = qa + (uint32_t)k;
                if (q >= p_hi) break;
                while (j < GPS_BIN_CHUNK - 1 && s_off[j + 1] <= q) ++j;
                const uint32_t slot = j0 + (uint32_t)j;
                if (slot >= ecnt) break;
                if (slot != cur) {
                    cur = slot;
                    const uint32_t r = list[e0 + slot];
                    pstart = coff[r] - pbias;
                    rec = load_record(recs, cidx[r]);
                }
                splat_point_region(rec, pstart + (q - s_off[j]), b, rx0, ry0, s_fb);
            }
        }
        __syncthreads();
    }
    // region -> framebuffer
    for (uint32_t w = (uint32_t)tid * 2u; w < words; w += 2u * GPS_BIN_THREADS) {
        const uint32_t t = w / tile_words, within = w - t * tile_words;
        const int tx = tx0 + (int)(t % (uint32_t)b.g.rtx), ty = ty0 + (int)(t / (uint32_t)b.g.rtx);
        if (tx >= b.g.tiles_per_row || ty >= tiles_rows) continue;
        unsigned long long* dst = fb + (uint64_t)(ty * b.g.tiles_per_row + tx) * tile_words + within;
        const unsigned long long a0 = s_fb[w], a1 = s_fb[w + 1];
        if (single) {
            *reinterpret_cast<uint4*>(dst) = make_uint4((uint32_t)a0, (uint32_t)(a0 >> 32), (uint32_t)a1,
                                                        (uint32_t)(a1 >> 32));
        } else {
            if (a0 != b.background) {
                atomicMin(dst, a0);
                ++merged;
            }
            if (a1 != b.background) {
                atomicMin(dst + 1, a1);
                ++merged;
            }
        }
    }
    __syncthreads();
    return merged;
}

// Persistent, one 1024-thread workgroup per CU (the region buffer is most of
// the LDS); jobs are taken from a device counter.
__global__ __launch_bounds__(GPS_BIN_THREADS) void bin_splat_kernel(
    const SplatRecord* __restrict__ recs, const uint32_t* __restrict__ cidx, const uint32_t* __restrict__ coff,
    const uint32_t* __restrict__ list, const uint32_t* __restrict__ lpre, const uint32_t* __restrict__ reg_ecnt,
    const uint32_t* __restrict__ reg_ptot, const uint32_t* __restrict__ reg_eoff, const uint32_t* __restrict__ reg_poff,
    const uint32_t* __restrict__ reg_jobs, const uint32_t* __restrict__ job_reg, const uint32_t* __restrict__ job_p0,
    uint32_t* __restrict__ state, unsigned long long* __restrict__ fb, BinSplat b, uint32_t job_points,
    uint32_t* __restrict__ atomics_count) {
    __shared__ unsigned long long s_fb[GPS_BIN_WORDS];
    __shared__ uint32_t s_off[GPS_BIN_CHUNK + 1];
    __shared__ uint32_t s_job, s_slot0;
    if (state[3]) return;
    const uint32_t njobs = state[1];
    uint32_t merged = 0;
    for (;;) {
        if (threadIdx.x == 0) s_job = atomicAdd(&state[4], 1u);
        __syncthreads();
        const uint32_t job = s_job;
        if (job >= njobs) break;
        const uint32_t reg = job_reg[job];
        const uint32_t p_lo = job_p0[job];
        const uint32_t p_hi = min(reg_ptot[reg], p_lo + job_points);
        merged += bin_splat_job(recs, cidx, coff, list, lpre, reg_eoff[reg], reg_ecnt[reg], reg_poff[reg], p_lo, p_hi,
                                reg_jobs[reg] == 1u, reg, b, fb, 0u, s_fb, s_off, &s_slot0);
    }
    if (atomics_count) add_count(atomics_count, merged);
}

// Gaussians too large to bin: all their points with global atomics, after
// the regions are written.
__global__ __launch_bounds__(256) void bin_big_kernel(const SplatRecord* __restrict__ recs,
                                                      const uint32_t* __restrict__ cidx,
                                                      const uint32_t* __restrict__ coff,
                                                      const uint32_t* __restrict__ ew, const uint32_t* __restrict__ big,
                                                      uint32_t* __restrict__ state, unsigned long long* __restrict__ fb,
                                                      SplatConsts sc, uint32_t* __restrict__ atomics_count) {
    __shared__ uint32_t s_e;
    if (state[3]) return;
    const int tid = threadIdx.x;
    const uint32_t nbig = state[2];
    uint32_t issued = 0;
    for (;;) {
        if (tid == 0) s_e = atomicAdd(&state[5], 1u);
        __syncthreads();
        const uint32_t e = s_e;
        __syncthreads();
        if (e >= nbig) break;
        const uint32_t r = big[e];
        const SplatRecord rec = load_record(recs, cidx[r]);
        const uint32_t p0 = coff[r], w = ew[r];
        int last = -1;  // one Gaussian throughout
        for (uint32_t k = (uint32_t)tid; k < w; k += 256u)
            issued += splat_point(rec, p0 + k, sc.frame_seed, sc.K, sc.spp, sc.up_w, sc.unbiased, fb, last);
    }
    if (atomics_count) add_count(atomics_count, issued);
}

// ------------------------------------------------------------------ binned splat, batches
// The regions of all binned views of a batch are numbered consecutively
// (view v owns [rbase(v), rbase(v) + nregions(v))), so the region scan, the
// list scans and the job queue are the single-frame kernels over the whole
// batch; counting and scattering take a grid row per view (grid.y = view)
// with the view's own histogram in LDS, and a job finds its view by its
// region.  Views whose bg.nregions is 0, or with more points than their
// dense_points, are left to the point-parallel batch splat (their regions stay
// empty).

__global__ __launch_bounds__(256) void batch_bin_count_kernel(const SplatRecord* __restrict__ recs,
                                                              const uint32_t* __restrict__ cidx,
                                                              const uint32_t* __restrict__ coff,
                                                              const BatchView* __restrict__ views,
                                                              uint2* __restrict__ erect, uint32_t* __restrict__ ew,
                                                              uint32_t* __restrict__ reg_cnt,
                                                              uint32_t* __restrict__ reg_pts,
                                                              uint32_t* __restrict__ big, uint32_t* __restrict__ state) {
    extern __shared__ uint32_t s_dyn[];  // 2 * the largest view's regions
    const BatchView& bv = views[blockIdx.y];
    const BinGeom g = bv.bg;
    if (g.nregions == 0u || bv.points > bv.dense_points) return;
    const uint32_t V = bv.valid, eb = bv.ebase;
    const uint32_t per = (V + gridDim.x - 1) / gridDim.x;
    const uint32_t r0 = eb + min(V, blockIdx.x * per), r1 = eb + min(V, blockIdx.x * per + per);
    bin_count_range(recs, cidx, coff, g, r0, r1, eb + V, bv.base + bv.total, bv.base + bv.points, bv.rbase, erect, ew,
                    reg_cnt, reg_pts, big, state, s_dyn, s_dyn + g.nregions);
}

__global__ __launch_bounds__(256) void batch_bin_scatter_kernel(const BatchView* __restrict__ views,
                                                                const uint2* __restrict__ erect,
                                                                const uint32_t* __restrict__ ew,
                                                                uint32_t* __restrict__ reg_cursor,
                                                                uint32_t* __restrict__ list, uint32_t* __restrict__ lw,
                                                                const uint32_t* __restrict__ state) {
    extern __shared__ uint32_t s_dyn[];
    if (state[3]) return;
    const BatchView& bv = views[blockIdx.y];
    const BinGeom g = bv.bg;
    if (g.nregions == 0u || bv.points > bv.dense_points) return;
    const uint32_t V = bv.valid, eb = bv.ebase;
    const uint32_t per = (V + gridDim.x - 1) / gridDim.x;
    const uint32_t r0 = eb + min(V, blockIdx.x * per), r1 = eb + min(V, blockIdx.x * per + per);
    bin_scatter_range(g, r0, r1, bv.rbase, erect, ew, reg_cursor, list, lw, s_dyn, s_dyn + g.nregions);
}

// The view that owns global region `reg`: the last view whose rbase <= reg
// (views without regions share the next binned view's rbase).
__device__ __forceinline__ int batch_region_view(const BatchView* __restrict__ views, int nviews, uint32_t reg) {
    int lo = 0, hi = nviews - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (views[mid].rbase <= reg) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(GPS_BIN_THREADS) void batch_bin_splat_kernel(
    const SplatRecord* __restrict__ recs, const uint32_t* __restrict__ cidx, const uint32_t* __restrict__ coff,
    const uint32_t* __restrict__ list, const uint32_t* __restrict__ lpre, const uint32_t* __restrict__ reg_ecnt,
    const uint32_t* __restrict__ reg_ptot, const uint32_t* __restrict__ reg_eoff, const uint32_t* __restrict__ reg_poff,
    const uint32_t* __restrict__ reg_jobs, const uint32_t* __restrict__ job_reg, const uint32_t* __restrict__ job_p0,
    uint32_t* __restrict__ state, const BatchView* __restrict__ views, int nviews,
    unsigned long long* __restrict__ fb, uint32_t job_points, uint32_t* __restrict__ atomics_count) {
    __shared__ unsigned long long s_fb[GPS_BIN_WORDS];
    __shared__ uint32_t s_off[GPS_BIN_CHUNK + 1];
    __shared__ uint32_t s_job, s_slot0;
    __shared__ int s_view;
    if (state[3]) return;
    const uint32_t njobs = state[1];
    uint32_t merged = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            s_job = atomicAdd(&state[4], 1u);
            s_view = s_job < njobs ? batch_region_view(views, nviews, job_reg[s_job]) : 0;
        }
        __syncthreads();
        const uint32_t job = s_job;
        if (job >= njobs) break;
        const BatchView& bv = views[s_view];
        BinSplat b;
        b.g = bv.bg;
        b.K = bv.f.Ki;
        b.unbiased = bv.f.unbiased;
        b.frame_seed = bv.f.frame_seed + 1u;
        b.background = bv.background;
        const uint32_t reg = job_reg[job];
        const uint32_t p_lo = job_p0[job];
        const uint32_t p_hi = min(reg_ptot[reg], p_lo + job_points);
        merged += bin_splat_job(recs, cidx, coff, list, lpre, reg_eoff[reg], reg_ecnt[reg], reg_poff[reg], p_lo, p_hi,
                                reg_jobs[reg] == 1u, reg - bv.rbase, b, fb + bv.fb_offset, bv.base, s_fb, s_off,
                                &s_slot0);
    }
    if (atomics_count) add_count(atomics_count, merged);
}

__global__ __launch_bounds__(256) void batch_bin_big_kernel(const SplatRecord* __restrict__ recs,
                                                            const uint32_t* __restrict__ cidx,
                                                            const uint32_t* __restrict__ coff,
                                                            const uint8_t* __restrict__ cview,
                                                            const uint32_t* __restrict__ ew,
                                                            const uint32_t* __restrict__ big,
                                                            uint32_t* __restrict__ state,
                                                            const BatchView* __restrict__ views,
                                                            unsigned long long* __restrict__ fb,
                                                            uint32_t* __restrict__ atomics_count) {
    __shared__ uint32_t s_e;
    if (state[3]) return;
    const int tid = threadIdx.x;
    const uint32_t nbig = state[2];
    uint32_t issued = 0;
    for (;;) {
        if (tid == 0) s_e = atomicAdd(&state[5], 1u);
        __syncthreads();
        const uint32_t e = s_e;
        __syncthreads();
        if (e >= nbig) break;
        const uint32_t r = big[e];
        const BatchView& bv = views[cview[r]];
        const SplatRecord rec = load_record(recs, cidx[r]);
        const uint32_t p0 = coff[r] - bv.base, w = ew[r];
        const uint32_t seed = bv.f.frame_seed + 1u;
        int last = -1;  // one Gaussian throughout
        for (uint32_t k = (uint32_t)tid; k < w; k += 256u)
            issued += splat_point(rec, p0 + k, seed, bv.f.Ki, bv.f.spp, bv.f.up_w, bv.f.unbiased, fb + bv.fb_offset,
                                  last);
    }
    if (atomics_count) add_count(atomics_count, issued);
}
// ------------------------------------------------------------------ combine
// One workgroup = 16x16 output pixels; lanes follow Morton order inside the
// block so a wave's framebuffer words are one contiguous range when ss is a
// power of two.
__global__ __launch_bounds__(256) void combine_kernel(const unsigned long long* __restrict__ fb, float* __restrict__ color,
                                                      float* __restrict__ depth, int W, int H, int ss, int spp,
                                                      int up_w, float near_plane, float far_plane, int depth_only) {
    const uint32_t t = threadIdx.x;
    uint32_t mx = 0, my = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        mx |= ((t >> (2 * b)) & 1u) << b;
        my |= ((t >> (2 * b + 1)) & 1u) << b;
    }
    const int x = blockIdx.x * 16 + (int)mx;
    const int y = blockIdx.y * 16 + (int)my;
    if (x >= W || y >= H) return;
    const int bx = x * ss, by = y * ss;
    float d = 1e9f;
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    for (int dx = 0; dx < ss; ++dx)
        for (int dy = 0; dy < ss; ++dy)
            for (int p = 0; p < spp; ++p) {
                const unsigned long long v = fb[fb_index(bx + dx, by + dy, p, spp, up_w)];
                const float vd = depth_dequantize((uint32_t)((v >> 36) & GPS_DEPTH_MAX), near_plane, far_plane);
                const float z = view_to_ndc_depth(vd, near_plane, far_plane);
                d = glm_min(d, z);
                if (depth_only) continue;
                cr += (float)((uint32_t)(v >> 24) & 0xFFFu) / 255.0f;
                cg += (float)((uint32_t)(v >> 12) & 0xFFFu) / 255.0f;
                cb += (float)((uint32_t)v & 0xFFFu) / 255.0f;
            }
    const uint32_t idx = (uint32_t)(y * W + x);
    if (!depth_only) {
        const float inv = 1.0f / (float)(spp * ss * ss);
        color[3 * idx] = cr * inv;
        color[3 * idx + 1] = cg * inv;
        color[3 * idx + 2] = cb * inv;
    }
    depth[idx] = d;
}

// Staged combine for power-of-two ss with ss^2 * spp <= 16: the framebuffer
// words of a 16x16 output block are one contiguous range (Morton order
// nests: morton(ss*x + dx, ss*y + dy) = ss^2 * morton(x, y) + morton(dx, dy)),
// so the workgroup streams it into LDS with 16-byte coalesced loads and
// every thread then reduces its own words in the reference's order
// (dx outer, dy, layer inner).  Pixel t's words sit at a stride of
// 2*wpp+1 dwords in LDS, which keeps the per-thread reads conflict-free.
template <int SS, int SPP>
__global__ __launch_bounds__(256) void combine_staged_kernel(const uint4* __restrict__ fb, float* __restrict__ color,
                                                             float* __restrict__ depth, int W, int H, int up_w,
                                                             float near_plane, float far_plane, int depth_only) {
    constexpr int WPP = SS * SS * SPP;          // words per output pixel
    constexpr int STRIDE = 2 * WPP + 1;         // dwords per pixel in LDS
    __shared__ uint32_t s[256 * STRIDE];
    const int t = threadIdx.x;
    // first FB pixel of this block and its word index
    const int fx0 = blockIdx.x * 16 * SS, fy0 = blockIdx.y * 16 * SS;
    const int64_t base_word = (int64_t)fb_index(fx0, fy0, 0, SPP, up_w);  // multiple of 256*WPP
    const uint4* src = fb + base_word / 2;
    constexpr int CHUNKS = 256 * WPP / 2;      // 16-byte chunks of the block
#pragma unroll
    for (int c = t; c < CHUNKS; c += 256) {
        const uint4 v = src[c];
        const int w0 = 2 * c;                  // first of the two words
        const int p0 = w0 / WPP, k0 = w0 % WPP;
        const int w1 = w0 + 1;
        const int p1 = w1 / WPP, k1 = w1 % WPP;
        s[p0 * STRIDE + 2 * k0] = v.x;
        s[p0 * STRIDE + 2 * k0 + 1] = v.y;
        s[p1 * STRIDE + 2 * k1] = v.z;
        s[p1 * STRIDE + 2 * k1 + 1] = v.w;
    }
    __syncthreads();
    uint32_t mx = 0, my = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        mx |= ((uint32_t)(t >> (2 * b)) & 1u) << b;
        my |= ((uint32_t)(t >> (2 * b + 1)) & 1u) << b;
    }
    const int x = blockIdx.x * 16 + (int)mx;
    const int y = blockIdx.y * 16 + (int)my;
    if (x >= W || y >= H) return;
    const uint32_t* mine = s + t * STRIDE;
    float d = 1e9f;
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
#pragma unroll
    for (int dx = 0; dx < SS; ++dx)
#pragma unroll
        for (int dy = 0; dy < SS; ++dy)
#pragma unroll
            for (int p = 0; p < SPP; ++p) {
                const int k = (int)morton_encode((uint32_t)dx, (uint32_t)dy) * SPP + p;
                const uint32_t lo = mine[2 * k], hi = mine[2 * k + 1];
                const float vd = depth_dequantize((hi >> 4) & (uint32_t)GPS_DEPTH_MAX, near_plane, far_plane);
                const float z = view_to_ndc_depth(vd, near_plane, far_plane);
                d = glm_min(d, z);
                if (depth_only) continue;
                cr += (float)(((lo >> 24) | (hi << 8)) & 0xFFFu) / 255.0f;
                cg += (float)((lo >> 12) & 0xFFFu) / 255.0f;
                cb += (float)(lo & 0xFFFu) / 255.0f;
            }
    const uint32_t idx = (uint32_t)(y * W + x);
    if (!depth_only) {
        const float inv = 1.0f / (float)(SPP * SS * SS);
        color[3 * idx] = cr * inv;
        color[3 * idx + 1] = cg * inv;
        color[3 * idx + 2] = cb * inv;
    }
    depth[idx] = d;
}

// ------------------------------------------------------------------ Hi-Z
// Each workgroup reduces a 32x32 block of the finest level through five
// levels in LDS; hiz_top_kernel then reduces the remaining coarse levels.
__global__ __launch_bounds__(256) void hiz_build_kernel(HiZView h, const float* __restrict__ depth, int kmin) {
    __shared__ float s0[1024], s1[1024];
    const int L = h.levels, fl = L - 1;
    const int tid = threadIdx.x;
    const int bx = blockIdx.x, by = blockIdx.y;
    float* cur = s0;
    float* nxt = s1;
    for (int i = tid; i < 1024; i += 256) {
        const int X = bx * 32 + (i & 31), Y = by * 32 + (i >> 5);
        float v = 0.0f;
        if (X < h.lw[fl] && Y < h.lh[fl]) {
            v = depth[Y * h.width + X];
            h.mip[h.off[fl] + (uint32_t)(Y * h.lw[fl] + X)] = v;
        }
        cur[i] = v;
    }
    __syncthreads();
    int size = 32;
    const int last_local = fl - 5 > 0 ? fl - 5 : 0;
    for (int l = fl - 1; l >= last_local; --l) {
        const int cs = size;
        size >>= 1;
        const bool use_min = ((L - 2) - l) < kmin;
        for (int i = tid; i < size * size; i += 256) {
            const int lx = i % size, ly = i / size;
            const int X = bx * size + lx, Y = by * size + ly;
            if (X < h.lw[l] && Y < h.lh[l]) {
                float v = use_min ? __builtin_bit_cast(float, 0x7F800000u) : __builtin_bit_cast(float, 0xFF800000u);
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        const int CX = 2 * X + dx, CY = 2 * Y + dy;
                        if (CX < h.lw[l + 1] && CY < h.lh[l + 1]) {
                            const float c = cur[(2 * ly + dy) * cs + (2 * lx + dx)];
                            v = use_min ? det_fmin(v, c) : det_fmax(v, c);
                        }
                    }
                h.mip[h.off[l] + (uint32_t)(Y * h.lw[l] + X)] = v;
                nxt[ly * size + lx] = v;
            }
        }
        __syncthreads();
        float* t = cur;
        cur = nxt;
        nxt = t;
    }
}

// Levels above the per-workgroup ones, in one workgroup (second launch; a
// device-wide ticket on one word serialized the 2000 workgroups).
// The level below `top` is staged in LDS once when it fits (up to 8192
// texels, frames to about 4K), so each higher level costs a barrier instead of
// a round trip through memory; larger frames read every level from memory.
__global__ __launch_bounds__(256) void hiz_top_kernel(HiZView h, int kmin, int top) {
    constexpr int CAP = 8192;
    __shared__ float s_a[CAP], s_b[CAP];
    const int L = h.levels;
    const int tid = threadIdx.x;
    const int n1 = h.lw[top + 1] * h.lh[top + 1];
    if (n1 <= CAP) {
        float* cur = s_a;
        float* nxt = s_b;
        for (int i = tid; i < n1; i += 256) cur[i] = h.mip[h.off[top + 1] + (uint32_t)i];
        __syncthreads();
        for (int l = top; l >= 0; --l) {
            const bool use_min = ((L - 2) - l) < kmin;
            const int n = h.lw[l] * h.lh[l];
            for (int i = tid; i < n; i += 256) {
                const int X = i % h.lw[l], Y = i / h.lw[l];
                float v = use_min ? __builtin_bit_cast(float, 0x7F800000u) : __builtin_bit_cast(float, 0xFF800000u);
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        const int CX = 2 * X + dx, CY = 2 * Y + dy;
                        if (CX < h.lw[l + 1] && CY < h.lh[l + 1]) {
                            const float c = cur[CY * h.lw[l + 1] + CX];
                            v = use_min ? det_fmin(v, c) : det_fmax(v, c);
                        }
                    }
                h.mip[h.off[l] + (uint32_t)i] = v;
                nxt[i] = v;
            }
            __syncthreads();
            float* t = cur;
            cur = nxt;
            nxt = t;
        }
        return;
    }
    for (int l = top; l >= 0; --l) {
        const bool use_min = ((L - 2) - l) < kmin;
        const int n = h.lw[l] * h.lh[l];
        for (int i = tid; i < n; i += 256) {
            const int X = i % h.lw[l], Y = i / h.lw[l];
            float v = use_min ? __builtin_bit_cast(float, 0x7F800000u) : __builtin_bit_cast(float, 0xFF800000u);
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const int CX = 2 * X + dx, CY = 2 * Y + dy;
                    if (CX < h.lw[l + 1] && CY < h.lh[l + 1]) {
                        const float c = h.mip[h.off[l + 1] + (uint32_t)(CY * h.lw[l + 1] + CX)];
                        v = use_min ? det_fmin(v, c) : det_fmax(v, c);
                    }
                }
            h.mip[h.off[l] + (uint32_t)i] = v;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(64) void zero_u32_kernel(uint32_t* __restrict__ p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += 64) p[i] = 0u;
}

__global__ __launch_bounds__(256) void fill_f32_kernel(float* __restrict__ p, uint32_t n, float v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------ culling blocks
// gaussian_bvh.cu:111-144 (one thread per block, Gaussians in index order).
template <int FORMAT>
__global__ __launch_bounds__(256) void build_nodes_kernel(DevScene s, uint32_t gpb, uint32_t nb, CullNode* __restrict__ nodes) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const uint32_t begin = b * gpb;
    const uint32_t end = (begin + gpb < s.N) ? begin + gpb : s.N;
    float mn[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
    float mx[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    for (uint32_t i = begin; i < end; ++i) {
        const Geo ge = load_geo<FORMAT>(s, i);
        float C[9];
        scale_rot_to_cov(ge.sx, ge.sy, ge.sz, ge.qw, ge.qx, ge.qy, ge.qz, C);
        const float ex = 3.0f * det_sqrt(C[0]), ey = 3.0f * det_sqrt(C[4]), ez = 3.0f * det_sqrt(C[8]);
        const float lo[3] = {ge.mx - ex, ge.my - ey, ge.mz - ez};
        const float hi[3] = {ge.mx + ex, ge.my + ey, ge.mz + ez};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            mn[k] = glm_min(mn[k], lo[k]);
            mx[k] = glm_max(mx[k], hi[k]);
        }
    }
    CullNode n;
    for (int k = 0; k < 3; ++k) {
        n.mn[k] = mn[k];
        n.mx[k] = mx[k];
    }
    n.begin = begin;
    n.count = end - begin;
    nodes[b] = n;
}

// ------------------------------------------------------------------ resolve
__global__ __launch_bounds__(256) void taa_kernel(const float* __restrict__ cur, const float* __restrict__ cur_depth,
                                                  const float* __restrict__ prev, Mat4 M, float blend, int W, int H,
                                                  int frame_number, float* __restrict__ out, uint8_t* __restrict__ rgb8) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= W * H) return;
    const int x = idx % W, y = idx / W;
    float c[3] = {cur[3 * idx], cur[3 * idx + 1], cur[3 * idx + 2]};
    const float depth = cur_depth[idx];
    if (frame_number > 0) {
        const float cx = ((float)x + 0.5f) / (float)W * 2.0f - 1.0f;
        const float cy = ((float)y + 0.5f) / (float)H * 2.0f - 1.0f;
        const float* m = M.m;
        const float px0 = (m[0] * cx + m[4] * cy) + (m[8] * depth + m[12] * 1.0f);
        const float py0 = (m[1] * cx + m[5] * cy) + (m[9] * depth + m[13] * 1.0f);
        const float pw0 = (m[3] * cx + m[7] * cy) + (m[11] * depth + m[15] * 1.0f);
        const float ppx = px0 / pw0, ppy = py0 / pw0;
        const float px = (ppx * 0.5f + 0.5f) * (float)W - 0.5f;
        const float py = (ppy * 0.5f + 0.5f) * (float)H - 0.5f;
        if (px >= 0.0f && px < (float)W && py >= 0.0f && py < (float)H) {
            const float flx = det_floor(px), fly = det_floor(py);
            int x0 = det_f2i_rz(flx), y0 = det_f2i_rz(fly);
            int x1 = x0 + 1, y1 = y0 + 1;
            x0 = (x0 < 0) ? 0 : ((x0 >= W) ? W - 1 : x0);
            y0 = (y0 < 0) ? 0 : ((y0 >= H) ? H - 1 : y0);
            x1 = (x1 < 0) ? 0 : ((x1 >= W) ? W - 1 : x1);
            y1 = (y1 < 0) ? 0 : ((y1 >= H) ? H - 1 : y1);
            const float tx = px - flx, ty = py - fly;
            const int i00 = y0 * W + x0, i10 = y0 * W + x1, i01 = y1 * W + x0, i11 = y1 * W + x1;
            float nmin[3] = {1e9f, 1e9f, 1e9f}, nmax[3] = {-1e9f, -1e9f, -1e9f};
            for (int dy = -1; dy <= 1; ++dy) {
                const int yy = y + dy;
                if (yy < 0 || yy >= H) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int xx = x + dx;
                    if (xx < 0 || xx >= W) continue;
                    const int j = yy * W + xx;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        nmin[k] = glm_min(nmin[k], cur[3 * j + k]);
                        nmax[k] = glm_max(nmax[k], cur[3 * j + k]);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float top = prev[3 * i00 + k] * (1.0f - tx) + prev[3 * i10 + k] * tx;
                const float bot = prev[3 * i01 + k] * (1.0f - tx) + prev[3 * i11 + k] * tx;
                const float hist = top * (1.0f - ty) + bot * ty;
                const float hc = glm_clamp(hist, nmin[k] - 0.05f, nmax[k] + 0.05f);
                c[k] = c[k] * (1.0f - blend) + hc * blend;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        out[3 * idx + k] = c[k];
        rgb8[3 * idx + k] = (uint8_t)__builtin_roundf(glm_clamp(c[k], 0.0f, 1.0f) * 255.0f);
    }
}

__global__ __launch_bounds__(256) void progressive_kernel(float* __restrict__ color, float* __restrict__ accum, int W,
                                                          int H, int n, uint8_t* __restrict__ rgb8) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= W * H) return;
    const float sc = 255.0f / (float)n;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float c = color[3 * idx + k];
        if (n > 1) c = c + accum[3 * idx + k];
        color[3 * idx + k] = c;
        accum[3 * idx + k] = 0.0f;
        rgb8[3 * idx + k] = (uint8_t)__builtin_roundf(glm_clamp(c * sc, 0.0f, 255.0f));
    }
}

// ------------------------------------------------------------------ debug
__global__ __launch_bounds__(256) void materialize_indices_kernel(const uint32_t* __restrict__ cidx,
                                                                  const uint32_t* __restrict__ coff,
                                                                  const PassCounters* __restrict__ ctr,
                                                                  uint32_t* __restrict__ out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ctr->points) return;
    uint32_t lo = 0, hi = ctr->valid - 1;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (coff[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    out[p] = cidx[lo];
}

// ------------------------------------------------------------------ multi-view batches
// A batch renders up to GPS_BATCH_MAX_VIEWS independent frames of one scene:
// view v is exactly the single-view frame (same seeds, same records, same
// framebuffer words), with its weights / records / tiles at offset v of
// concatenated arrays, so the distributor and a persistent splat cover all
// views in one launch each.

__global__ __launch_bounds__(256) void batch_clear_kernel(const BatchView* __restrict__ views, uint4* __restrict__ fb) {
    const BatchView& bv = views[blockIdx.y];
    const uint64_t n2 = bv.fb_words / 2;
    const uint64_t value = bv.background;
    const uint4 v = make_uint4((uint32_t)value, (uint32_t)(value >> 32), (uint32_t)value, (uint32_t)(value >> 32));
    uint4* dst = fb + bv.fb_offset / 2;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

// Frustum part of the block cull (gaussian_bvh.cu:215-234) per view.
__global__ __launch_bounds__(256) void batch_cull_kernel(BatchView* __restrict__ views, CullView cv,
                                                         uint32_t mask_stride) {
    const uint32_t v = blockIdx.y;
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const float* pl = views[v].planes;
    bool culled = false;
    if (idx < cv.node_count && views[v].f.hier) {  // views without block culling get an empty mask
        const CullNode n = cv.nodes[idx];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const float nx = pl[4 * i], ny = pl[4 * i + 1], nz = pl[4 * i + 2], d = pl[4 * i + 3];
            const float vx = nx >= 0.0f ? n.mx[0] : n.mn[0];
            const float vy = ny >= 0.0f ? n.mx[1] : n.mn[1];
            const float vz = nz >= 0.0f ? n.mx[2] : n.mn[2];
            if (((nx * vx + ny * vy) + nz * vz) + d < 0.0f) culled = true;
        }
    }
    const uint64_t b = __ballot(culled);
    const int lane = threadIdx.x & 63;
    if (lane == 0) {
        uint32_t* mask = cv.mask + (uint64_t)v * mask_stride;
        const uint32_t w0 = (idx - lane) >> 5;
        if (w0 < cv.mask_words) mask[w0] = (uint32_t)b;
        if (w0 + 1 < cv.mask_words) mask[w0 + 1] = (uint32_t)(b >> 32);
    }
    add_count(&views[v].culled_blocks, culled ? 1u : 0u);  // the workgroup's view: one atomic per workgroup
}

// One workgroup per tile of 4096 Gaussians; each thread evaluates its 16
// Gaussians under every view (view loop inner, so a Gaussian's attributes
// are fetched from HBM once per batch and from the caches afterwards).  The
// per-(view, tile) sums are complete when the workgroup ends, so they are
// stored, not accumulated.
template <int FORMAT, int DEG>
__global__ __launch_bounds__(256) void batch_preprocess_kernel(DevScene s, const BatchView* __restrict__ views,
                                                               int nviews, int views_per_wg, uint32_t npad,
                                                               uint32_t tiles_per_view,
                                                               uint32_t* __restrict__ weights,
                                                               SplatRecord* __restrict__ recs,
                                                               uint32_t* __restrict__ tile_sum,
                                                               uint32_t* __restrict__ tile_cnt, CullView cv,
                                                               uint32_t mask_stride) {
    __shared__ uint32_t s_sum[GPS_BATCH_MAX_VIEWS], s_cnt[GPS_BATCH_MAX_VIEWS];
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t tile = blockIdx.x;
    // grid row y evaluates views [v0, v1): the (view, tile) sums of a view
    // come from one workgroup whatever the split
    const int v0 = (int)blockIdx.y * views_per_wg, v1 = min(nviews, v0 + views_per_wg);
    for (int v = v0 + tid; v < v1; v += 256) {
        s_sum[v] = 0u;
        s_cnt[v] = 0u;
    }
    __syncthreads();
    const HiZView no_hiz{};
    for (int i = 0; i < GPS_TILE / 256; ++i) {
        const uint32_t g = (tile << GPS_TILE_SHIFT) + (uint32_t)(i * 256 + tid);
        for (int v = v0; v < v1; ++v) {
            const FrameConsts& f = views[v].f;
            uint32_t n = 0;
            if (g < s.N) {
                const uint32_t blk = g / cv.gpb;
                const bool culled = f.hier && mask_bit(cv.mask + (uint64_t)v * mask_stride, blk);
                if (!culled) {
                    n = preprocess_one<FORMAT, DEG>(s, f, g, nullptr, no_hiz, recs + (uint64_t)v * npad + g);
                    weights[(uint64_t)v * npad + g] = n;
                }
            }
            const uint64_t b = __ballot(n != 0u);
            if (b) {
                const uint32_t sum = wave_reduce_add(n);
                if (lane == 0) {
                    atomicAdd(&s_sum[v], sum);
                    atomicAdd(&s_cnt[v], (uint32_t)__popcll(b));
                }
            }
        }
    }
    __syncthreads();
    for (int v = v0 + tid; v < v1; v += 256) {
        tile_sum[(uint64_t)v * tiles_per_view + tile] = s_sum[v];
        tile_cnt[(uint64_t)v * tiles_per_view + tile] = s_cnt[v];
    }
}

// Per-view point ranges of the batch from the concatenated tile scan.
__global__ __launch_bounds__(256) void batch_views_kernel(BatchView* __restrict__ views, int nviews,
                                                          const uint32_t* __restrict__ tile_off,
                                                          const uint32_t* __restrict__ tile_voff,
                                                          uint32_t tiles_per_view, const PassCounters* __restrict__ ctr) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nviews) return;
    const uint64_t t0 = (uint64_t)v * tiles_per_view, t1 = t0 + tiles_per_view;
    const uint32_t base = tile_off[t0];
    const uint32_t end = v + 1 < nviews ? tile_off[t1] : ctr->total;
    const uint32_t vbase = tile_voff[t0];
    const uint32_t vend = v + 1 < nviews ? tile_voff[t1] : ctr->valid;
    views[v].base = base;
    views[v].total = end - base;
    views[v].points = min(end - base, views[v].budget);
    views[v].valid = vend - vbase;
    views[v].ebase = vbase;
}

// Persistent splat over the points of all views (same schedule as
// splat_kernel, THREADS x PPT points per chunk).  Point p of view v is point
// p - base(v) of that view's own frame.  Views with binned regions are
// skipped unless the binned kernels stood down (bin_state[3], list overflow)
// or the view is too dense to bin (points > dense_points).
template <int THREADS, int PPT>
__global__ __launch_bounds__(THREADS) void batch_splat_kernel(const SplatRecord* __restrict__ recs,
                                                              const uint32_t* __restrict__ cidx,
                                                              const uint32_t* __restrict__ coff,
                                                              const uint8_t* __restrict__ cview,
                                                              const uint32_t* __restrict__ chunk_start,
                                                              const PassCounters* __restrict__ ctr,
                                                              const BatchView* __restrict__ views,
                                                              unsigned long long* __restrict__ fb, int32_t xcds,
                                                              uint32_t* __restrict__ atomics_count,
                                                              const uint32_t* __restrict__ bin_state,
                                                              int use_table) {
    constexpr int CHUNK = THREADS * PPT;
    const bool bin_ran = bin_state && bin_state[3] == 0u;
    __shared__ uint32_t s_off[CHUNK + 1];
#if GPS_SPLAT_SLOTS
    // word table as in splat_kernel, keyed by the word's index in the whole
    // batch framebuffer (use_table: that index fits 32 bits)
    __shared__ uint32_t s_key_[GPS_SPLAT_SLOTS];
    __shared__ unsigned long long s_val_[GPS_SPLAT_SLOTS];
    uint32_t* const s_key = use_table ? s_key_ : nullptr;
    unsigned long long* const s_val = s_val_;
#else
    uint32_t* const s_key = nullptr;
    unsigned long long* const s_val = nullptr;
#endif
    const uint32_t P = ctr->points;
    const uint32_t V = ctr->valid;
    const uint32_t nchunks = (uint32_t)(((uint64_t)P + CHUNK - 1) / CHUNK);
    const int tid = threadIdx.x;
    uint32_t issued = 0;
#if GPS_SPLAT_SLOTS
    if (s_key)
        for (int i = tid; i < GPS_SPLAT_SLOTS; i += THREADS) {
            s_key[i] = 0xFFFFFFFFu;
            s_val[i] = ~0ull;
        }
#endif
    const uint32_t nx = (uint32_t)xcds;
    const uint32_t xcd = blockIdx.x % nx;
    const uint32_t per_xcd = gridDim.x / nx;
    const uint32_t seg_lo = (uint32_t)(((uint64_t)nchunks * xcd) / nx);
    const uint32_t seg_hi = (uint32_t)(((uint64_t)nchunks * (xcd + 1)) / nx);
    for (uint32_t c = seg_lo + blockIdx.x / nx; c < seg_hi; c += per_xcd) {
        const uint32_t r_lo = chunk_start[c];
        for (int j = tid; j <= CHUNK; j += THREADS) {
            const uint32_t r = r_lo + (uint32_t)j;
            s_off[j] = r < V ? coff[r] : 0xFFFFFFFFu;
        }
        __syncthreads();
        const uint32_t p0 = c * CHUNK + (uint32_t)tid * PPT;
        if (p0 < P) {
            int lo = 0, hi = CHUNK;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= p0) lo = mid;
                else hi = mid - 1;
            }
            int j = lo;
            uint32_t cur_g = 0xFFFFFFFFu;
            int cur_v = -1;
            SplatRecord rec;
            uint32_t vbase = 0, vpoints = 0, vseed = 0;
            int vK = 1, vspp = 1, vup_w = 1, vunb = 1;
            bool vskip = false;
            uint64_t voff = 0;
            int last = -1;
            for (int k = 0; k < PPT; ++k) {
                const uint32_t p = p0 + (uint32_t)k;
                if (p >= P) break;
                while (j < CHUNK - 1 && s_off[j + 1] <= p) ++j;
                if (r_lo + (uint32_t)j >= V) break;
                const uint32_t r = r_lo + (uint32_t)j;
                const uint32_t g = cidx[r];
                if (g != cur_g) {
                    cur_g = g;
                    const int v = cview[r];
                    if (v != cur_v) {
                        cur_v = v;
                        const BatchView& bv = views[v];
                        vbase = bv.base;
                        vpoints = bv.points;
                        vseed = bv.f.frame_seed + 1u;
                        vK = bv.f.Ki;
                        vspp = bv.f.spp;
                        vup_w = bv.f.up_w;
                        vunb = bv.f.unbiased;
                        voff = bv.fb_offset;
                        vskip = bin_ran && bv.bg.nregions != 0u && bv.points <= bv.dense_points;
                    }
                    if (!vskip) rec = load_record(recs, g);
                    last = -1;
                }
                if (vskip) continue;  // rendered by the binned kernels
                const uint32_t pl = p - vbase;
                if (pl >= vpoints) continue;  // beyond the view's point budget
                issued += splat_point(rec, pl, vseed, vK, vspp, vup_w, vunb, fb + voff, last, s_key, s_val,
                                      (uint32_t)voff);
            }
        }
        __syncthreads();
#if GPS_SPLAT_SLOTS
        if (s_key) issued += flush_word_table<THREADS>(s_key, s_val, fb);
#endif
    }
    if (atomics_count) add_count(atomics_count, issued);
}

// ------------------------------------------------------------------ launchers
template <int FORMAT, int DEG>
static void launch_pre(const PreprocessArgs& a, hipStream_t st) {
    const uint32_t blocks = (a.scene.N + GPS_PRE_THREADS - 1) / GPS_PRE_THREADS;
    if (a.rejects)  // reject-reason counters: the work-list kernel (same outputs) with counting
        hipLaunchKernelGGL((preprocess_kernel<FORMAT, DEG, true>), dim3(blocks), dim3(GPS_PRE_THREADS), 0, st,
                           a.scene, a.f, a.maskA, a.maskB, a.iteration, a.weights, a.recs, a.tile_sum, a.tile_cnt,
                           a.cull, a.hiz, a.gocc, a.rejects, a.wg_live);
    else if (a.f.hier || a.f.occl)  // culled frames: work lists keep the later phases on full waves
        hipLaunchKernelGGL((preprocess_kernel<FORMAT, DEG, false>), dim3(blocks), dim3(GPS_PRE_THREADS), 0, st,
                           a.scene, a.f, a.maskA, a.maskB, a.iteration, a.weights, a.recs, a.tile_sum, a.tile_cnt,
                           a.cull, a.hiz, a.gocc, nullptr, a.wg_live);
    else
        hipLaunchKernelGGL((preprocess_flat_kernel<FORMAT, DEG>), dim3(blocks), dim3(GPS_PRE_THREADS), 0, st,
                           a.scene, a.f, a.maskA, a.maskB, a.iteration, a.weights, a.recs, a.tile_sum, a.tile_cnt,
                           a.cull, a.hiz);
}

void launch_preprocess(const PreprocessArgs& a, hipStream_t st) {
    if (a.scene.N == 0) return;
    const int d = a.scene.sh_degree;
    if (a.scene.format == GPS_FORMAT_FP32) {
        if (d == 0) launch_pre<GPS_FORMAT_FP32, 0>(a, st);
        else if (d == 1) launch_pre<GPS_FORMAT_FP32, 1>(a, st);
        else if (d == 2) launch_pre<GPS_FORMAT_FP32, 2>(a, st);
        else launch_pre<GPS_FORMAT_FP32, 3>(a, st);
    } else if (a.scene.format == GPS_FORMAT_QUANTIZED) {
        launch_pre<GPS_FORMAT_QUANTIZED, 3>(a, st);  // the Sloan basis always has 16 coefficients
    } else {
        if (d == 0) launch_pre<GPS_FORMAT_HALF_FP8, 0>(a, st);
        else if (d == 1) launch_pre<GPS_FORMAT_HALF_FP8, 1>(a, st);
        else if (d == 2) launch_pre<GPS_FORMAT_HALF_FP8, 2>(a, st);
        else launch_pre<GPS_FORMAT_HALF_FP8, 3>(a, st);
    }
}

void launch_clear(uint64_t* fb, uint64_t words, uint64_t value, hipStream_t st, uint32_t* z0, uint32_t n0,
                  uint32_t* z1, uint32_t n1) {
    const uint64_t n2 = words / 2;
    uint64_t blocks = (n2 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0 && n0 + n1 == 0) return;
    if (n0 > 256 || n1 > 256) {  // more than one workgroup zeroes: separate launches
        if (n0) launch_zero_u32(z0, n0, st);
        if (n1) launch_zero_u32(z1, n1, st);
        n0 = n1 = 0;
        if (blocks == 0) return;
    }
    hipLaunchKernelGGL(clear_kernel, dim3(blocks ? (uint32_t)blocks : 1u), dim3(256), 0, st,
                       reinterpret_cast<uint4*>(fb), n2, value, z0, n0, z1, n1);
}

void launch_cull(const CullView& cv, const float planes[24], const Mat4& vp, const HiZView& h, int occl,
                 uint32_t* culled_count, uint32_t* live, uint32_t* live_count, uint32_t* maskA, uint32_t* maskB,
                 const uint32_t* iteration, uint8_t* wg_live, uint32_t N, hipStream_t st) {
    if (cv.node_count == 0) return;
    Planes pl;
    for (int i = 0; i < 24; ++i) pl.p[i] = planes[i];
    const uint32_t blocks = (cv.node_count + 255) / 256;
    const MaskZero mz{maskA, maskB, iteration, wg_live, N};
    hipLaunchKernelGGL(cull_kernel, dim3(blocks), dim3(256), 0, st, cv, pl, vp, h, occl, culled_count, live,
                       live_count, mz);
}

// One tile-scan workgroup per 8192 tiles, at least one (it publishes the
// pass totals also when there are no tiles).
static uint32_t tile_scan_grid(uint32_t ntiles) { return ntiles ? (ntiles + 8191u) / 8192u : 1u; }

void launch_distribute(const DistributeArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(tile_scan_kernel, dim3(tile_scan_grid(a.ntiles)), dim3(1024), 0, st, a.tile_sum, a.tile_cnt,
                       a.ntiles, a.tile_off, a.tile_voff, a.ctr, a.budget, a.iteration);
    if (a.ntiles == 0) return;
    hipLaunchKernelGGL(distribute_kernel, dim3(a.ntiles), dim3(256), 0, st, a.weights, a.tile_off, a.tile_voff,
                       a.ntiles, a.ctr, a.cidx, a.coff, a.chunk_start, a.cull, a.hier, a.N, a.ntiles, 0u, 0u,
                       (uint8_t*)nullptr, GPS_SPLAT_CHUNK_SHIFT, a.tile_sum, a.tile_cnt);
}

void launch_ordered_splat(const OrderedArgs& a, hipStream_t st) {
    const SplatArgs& s = a.s;
    SplatConsts sc;
    sc.spp = s.spp;
    sc.K = s.K;
    sc.up_w = s.up_w;
    sc.unbiased = s.unbiased;
    sc.frame_seed = s.frame_seed;
    sc.xcds = s.xcds;
    OrdGate gate;
    gate.K = (uint32_t)s.K;
    gate.up_w = (uint32_t)a.up_w;
    gate.up_h = (uint32_t)a.up_h;
    gate.forced = a.forced;
    unsigned long long* fb = reinterpret_cast<unsigned long long*>(s.fb);
    hipLaunchKernelGGL(ord_key_kernel, dim3(a.cus), dim3(256), 0, st, s.recs, s.cidx, s.chunk_start, s.ctr, gate,
                       s.only_if, a.bucket, a.hist);
    hipLaunchKernelGGL(ord_scan_kernel, dim3(1), dim3(1024), 0, st, s.ctr, gate, s.only_if, a.hist, a.lim);
    for (int ph = 0; ph < GPS_ORD_PHASES; ++ph) {
        const bool test = ph > 0 || a.pre_test;
        if (test)
            hipLaunchKernelGGL(ord_tile_max_kernel, dim3((a.ntiles + 3u) / 4u), dim3(256), 0, st,
                               (const unsigned long long*)fb, a.tmax, a.ntiles, s.spp, s.ctr, gate, s.only_if);
        hipLaunchKernelGGL(splat_ordered_kernel, dim3(s.grid), dim3(GPS_SPLAT_THREADS), 0, st, s.recs, s.cidx, s.coff,
                           s.chunk_start, s.ctr, fb, sc, s.atomics_count, s.only_if, (const uint16_t*)a.bucket,
                           (const uint32_t*)a.lim, (const uint32_t*)a.tmax, a.tw64, ph, test ? 1 : 0, gate,
                           a.skipped_count);
    }
}

void launch_splat(const SplatArgs& a, hipStream_t st) {
    SplatConsts sc;
    sc.spp = a.spp;
    sc.K = a.K;
    sc.up_w = a.up_w;
    sc.unbiased = a.unbiased;
    sc.frame_seed = a.frame_seed;
    sc.xcds = a.xcds;
    hipLaunchKernelGGL(splat_kernel, dim3(a.grid), dim3(GPS_SPLAT_THREADS), 0, st, a.recs, a.cidx, a.coff,
                       a.chunk_start, a.ctr, reinterpret_cast<unsigned long long*>(a.fb), sc, a.atomics_count,
                       a.only_if);
}

template <int SS, int SPP>
static bool try_staged(const uint64_t* fb, float* color, float* depth, int W, int H, int ss, int spp, float near_plane,
                       float far_plane, int depth_only, hipStream_t st) {
    if (ss != SS || spp != SPP) return false;
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    hipLaunchKernelGGL((combine_staged_kernel<SS, SPP>), grid, dim3(256), 0, st, reinterpret_cast<const uint4*>(fb),
                       color, depth, W, H, W * ss, near_plane, far_plane, depth_only);
    return true;
}

void launch_combine(const uint64_t* fb, float* color, float* depth, int W, int H, int ss, int spp, float near_plane,
                    float far_plane, int depth_only, hipStream_t st) {
    if (try_staged<1, 1>(fb, color, depth, W, H, ss, spp, near_plane, far_plane, depth_only, st) ||
        try_staged<1, 2>(fb, color, depth, W, H, ss, spp, near_plane, far_plane, depth_only, st) ||
        try_staged<1, 4>(fb, color, depth, W, H, ss, spp, near_plane, far_plane, depth_only, st) ||
        try_staged<2, 1>(fb, color, depth, W, H, ss, spp, near_plane, far_plane, depth_only, st) ||
        try_staged<2, 2>(fb, color, depth, W, H, ss, spp, near_plane, far_plane, depth_only, st) ||
        try_staged<2, 4>(fb, color, depth, W, H, ss, spp, near_plane, far_plane, depth_only, st) ||
        try_staged<4, 1>(fb, color, depth, W, H, ss, spp, near_plane, far_plane, depth_only, st))
        return;
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    hipLaunchKernelGGL(combine_kernel, grid, dim3(256), 0, st, reinterpret_cast<const unsigned long long*>(fb), color,
                       depth, W, H, ss, spp, W * ss, near_plane, far_plane, depth_only);
}

void launch_hiz_build(const HiZView& h, const float* depth, int kmin, uint32_t* ticket, hipStream_t st) {
    dim3 grid((h.width + 31) / 32, (h.height + 31) / 32);
    (void)ticket;
    hipLaunchKernelGGL(hiz_build_kernel, grid, dim3(256), 0, st, h, depth, kmin);
    const int fl = h.levels - 1, last_local = fl - 5 > 0 ? fl - 5 : 0;
    if (last_local > 0) hipLaunchKernelGGL(hiz_top_kernel, dim3(1), dim3(256), 0, st, h, kmin, last_local - 1);
}

void launch_fill_f32(float* p, uint32_t n, float v, hipStream_t st) {
    uint32_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) return;
    hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks), dim3(256), 0, st, p, n, v);
}

void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, st, p, n);
}

void launch_group_boxes(const DevScene& s, GroupBox* boxes, uint32_t count, hipStream_t st) {
    if (count == 0) return;
    const uint32_t blocks = (count + 3) / 4;  // four groups (waves) per workgroup
    if (s.format == GPS_FORMAT_FP32)
        hipLaunchKernelGGL(group_box_kernel<GPS_FORMAT_FP32>, dim3(blocks), dim3(256), 0, st, s, boxes, count);
    else if (s.format == GPS_FORMAT_QUANTIZED)
        hipLaunchKernelGGL(group_box_kernel<GPS_FORMAT_QUANTIZED>, dim3(blocks), dim3(256), 0, st, s, boxes, count);
    else
        hipLaunchKernelGGL(group_box_kernel<GPS_FORMAT_HALF_FP8>, dim3(blocks), dim3(256), 0, st, s, boxes, count);
}

void launch_group_cull(const GroupView& gv, const FrameConsts& f, const HiZView& h, uint32_t* culled_count,
                       hipStream_t st) {
    if (gv.count == 0) return;
    hipLaunchKernelGGL(group_cull_kernel, dim3((gv.count + 255) / 256), dim3(256), 0, st, gv, f, h, culled_count);
}

void launch_group_cull_live(const GroupView& gv, const CullView& cv, const uint32_t* live, const uint32_t* live_count,
                            uint32_t N, const FrameConsts& f, const HiZView& h, uint32_t grid,
                            uint32_t* culled_count, hipStream_t st) {
    if (gv.count == 0 || cv.node_count == 0) return;
    // groups a block meets: gpb / 64 when blocks start on group boundaries, else one more
    const uint32_t per_block = (cv.gpb % GPS_GROUP == 0) ? cv.gpb / GPS_GROUP : (cv.gpb + GPS_GROUP - 1) / GPS_GROUP + 1;
    hipLaunchKernelGGL(group_cull_live_kernel, dim3(grid), dim3(256), 0, st, gv, cv, live, live_count, per_block, N, f,
                       h, culled_count);
}

void launch_build_nodes(const DevScene& s, uint32_t gpb, uint32_t nb, CullNode* nodes, hipStream_t st) {
    if (nb == 0) return;
    const uint32_t blocks = (nb + 255) / 256;
    if (s.format == GPS_FORMAT_FP32)
        hipLaunchKernelGGL(build_nodes_kernel<GPS_FORMAT_FP32>, dim3(blocks), dim3(256), 0, st, s, gpb, nb, nodes);
    else if (s.format == GPS_FORMAT_QUANTIZED)
        hipLaunchKernelGGL(build_nodes_kernel<GPS_FORMAT_QUANTIZED>, dim3(blocks), dim3(256), 0, st, s, gpb, nb,
                           nodes);
    else
        hipLaunchKernelGGL(build_nodes_kernel<GPS_FORMAT_HALF_FP8>, dim3(blocks), dim3(256), 0, st, s, gpb, nb, nodes);
}

void launch_taa(const float* cur, const float* cur_depth, const float* prev, const Mat4& M, float blend, int W, int H,
                int frame_number, float* out, uint8_t* rgb8, hipStream_t st) {
    const int n = W * H;
    hipLaunchKernelGGL(taa_kernel, dim3((n + 255) / 256), dim3(256), 0, st, cur, cur_depth, prev, M, blend, W, H,
                       frame_number, out, rgb8);
}

void launch_progressive(float* color, float* accum, int W, int H, int n, uint8_t* rgb8, hipStream_t st) {
    const int cnt = W * H;
    hipLaunchKernelGGL(progressive_kernel, dim3((cnt + 255) / 256), dim3(256), 0, st, color, accum, W, H, n, rgb8);
}

void launch_materialize_indices(const uint32_t* cidx, const uint32_t* coff, const PassCounters* ctr, uint32_t* out,
                                uint32_t max_points, hipStream_t st) {
    if (max_points == 0) return;
    hipLaunchKernelGGL(materialize_indices_kernel, dim3((max_points + 255) / 256), dim3(256), 0, st, cidx, coff, ctr, out);
}

// ------------------------------------------------------------------ batch launchers
void launch_batch_clear(const BatchArgs& a, hipStream_t st) {
    uint32_t bx = (uint32_t)std::min<uint64_t>((a.max_fb_words / 2 + 255) / 256, 2048);
    if (bx == 0) bx = 1;
    hipLaunchKernelGGL(batch_clear_kernel, dim3(bx, a.nviews), dim3(256), 0, st, a.views,
                       reinterpret_cast<uint4*>(a.fb));
}

void launch_batch_cull(const BatchArgs& a, hipStream_t st) {
    if (!a.cull.node_count) return;
    hipLaunchKernelGGL(batch_cull_kernel, dim3((a.cull.node_count + 255) / 256, a.nviews), dim3(256), 0, st, a.views,
                       a.cull, a.mask_stride);
}

template <int FORMAT, int DEG> static void launch_batch_pre_t(const BatchArgs& a, hipStream_t st) {
    const int vpw = a.pre_views > 0 && a.pre_views < a.nviews ? a.pre_views : a.nviews;
    hipLaunchKernelGGL((batch_preprocess_kernel<FORMAT, DEG>), dim3(a.tiles_per_view, (a.nviews + vpw - 1) / vpw),
                       dim3(256), 0, st, a.scene, a.views, a.nviews, vpw, a.npad, a.tiles_per_view, a.weights, a.recs,
                       a.tile_sum, a.tile_cnt, a.cull, a.mask_stride);
}

void launch_batch_preprocess(const BatchArgs& a, hipStream_t st) {
    if (a.scene.N == 0) return;
    const int d = a.scene.sh_degree;
    if (a.scene.format == GPS_FORMAT_FP32) {
        if (d == 0) launch_batch_pre_t<GPS_FORMAT_FP32, 0>(a, st);
        else if (d == 1) launch_batch_pre_t<GPS_FORMAT_FP32, 1>(a, st);
        else if (d == 2) launch_batch_pre_t<GPS_FORMAT_FP32, 2>(a, st);
        else launch_batch_pre_t<GPS_FORMAT_FP32, 3>(a, st);
    } else if (a.scene.format == GPS_FORMAT_QUANTIZED) {
        launch_batch_pre_t<GPS_FORMAT_QUANTIZED, 3>(a, st);
    } else {
        if (d == 0) launch_batch_pre_t<GPS_FORMAT_HALF_FP8, 0>(a, st);
        else if (d == 1) launch_batch_pre_t<GPS_FORMAT_HALF_FP8, 1>(a, st);
        else if (d == 2) launch_batch_pre_t<GPS_FORMAT_HALF_FP8, 2>(a, st);
        else launch_batch_pre_t<GPS_FORMAT_HALF_FP8, 3>(a, st);
    }
}

void launch_batch_distribute(const BatchArgs& a, hipStream_t st) {
    const uint32_t ntiles = a.tiles_per_view * (uint32_t)a.nviews;
    hipLaunchKernelGGL(tile_scan_kernel, dim3(tile_scan_grid(ntiles)), dim3(1024), 0, st, a.tile_sum, a.tile_cnt,
                       ntiles, a.tile_off, a.tile_voff, a.ctr, 0xFFFFFFFFu, a.iteration);
    hipLaunchKernelGGL(batch_views_kernel, dim3((a.nviews + 255) / 256), dim3(256), 0, st, a.views, a.nviews,
                       (const uint32_t*)a.tile_off, (const uint32_t*)a.tile_voff, a.tiles_per_view,
                       (const PassCounters*)a.ctr);
    if (ntiles == 0) return;
    uint32_t shift = 0;
    while ((1u << shift) < (uint32_t)(a.threads * a.ppt)) ++shift;
    hipLaunchKernelGGL(distribute_kernel, dim3(ntiles), dim3(256), 0, st, (const uint32_t*)a.weights,
                       (const uint32_t*)a.tile_off, (const uint32_t*)a.tile_voff, ntiles, (const PassCounters*)a.ctr,
                       a.cidx, a.coff, a.chunk_start, a.cull, a.hier, a.scene.N, a.tiles_per_view, a.npad,
                       a.mask_stride, a.cview, shift, a.tile_sum, a.tile_cnt);
}

template <int THREADS, int PPT> static void launch_batch_splat_t(const BatchArgs& a, hipStream_t st) {
    hipLaunchKernelGGL((batch_splat_kernel<THREADS, PPT>), dim3(a.grid), dim3(THREADS), 0, st,
                       (const SplatRecord*)a.recs, (const uint32_t*)a.cidx, (const uint32_t*)a.coff,
                       (const uint8_t*)a.cview, (const uint32_t*)a.chunk_start, (const PassCounters*)a.ctr,
                       (const BatchView*)a.views, reinterpret_cast<unsigned long long*>(a.fb), a.xcds, a.atomics,
                       a.bin.any ? (const uint32_t*)a.bin.state : nullptr, a.fb_key32);
}

// Splat variants the batch autotuner chooses from: (threads, points per thread).
const int kBatchSplatVariants[][2] = {{256, 4}, {256, 2}, {256, 8}, {128, 4}, {512, 2}, {512, 4}};
const int kNumBatchSplatVariants = 6;
// Preprocess splits the batch autotuner chooses from: views per workgroup
// (0: all views of the batch in one workgroup per tile).
const int kBatchPreViews[] = {0, 16, 4};
const int kNumBatchPreViews = 3;

bool launch_batch_splat(const BatchArgs& a, hipStream_t st) {
    const int t = a.threads, q = a.ppt;
    if (t == 256 && q == 4) launch_batch_splat_t<256, 4>(a, st);
    else if (t == 256 && q == 2) launch_batch_splat_t<256, 2>(a, st);
    else if (t == 256 && q == 8) launch_batch_splat_t<256, 8>(a, st);
    else if (t == 128 && q == 4) launch_batch_splat_t<128, 4>(a, st);
    else if (t == 512 && q == 2) launch_batch_splat_t<512, 2>(a, st);
    else if (t == 512 && q == 4) launch_batch_splat_t<512, 4>(a, st);
    else return false;
    return true;
}

void launch_batch_binned(const BatchArgs& a, hipStream_t st) {
    const BatchBinArgs& b = a.bin;
    if (!b.any) return;
    launch_zero_u32(b.state, 8, st);
    uint32_t gx = (b.cus * 4u + (uint32_t)a.nviews - 1u) / (uint32_t)a.nviews;  // workgroups per view
    if (gx < 2u) gx = 2u;
    const size_t hist = 2u * b.max_view_regions * sizeof(uint32_t);
    hipLaunchKernelGGL(batch_bin_count_kernel, dim3(gx, a.nviews), dim3(256), hist, st, (const SplatRecord*)a.recs,
                       (const uint32_t*)a.cidx, (const uint32_t*)a.coff, (const BatchView*)a.views, b.erect, b.ew,
                       b.reg_cnt, b.reg_pts, b.big, b.state);
    BinGeom all{};
    all.nregions = b.nregions;
    hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, st, all, b.reg_cnt, b.reg_pts, b.reg_ecnt, b.reg_ptot,
                       b.reg_eoff, b.reg_poff, b.reg_cursor, b.reg_jobs, b.job_reg, b.job_p0, b.state, b.list_cap,
                       b.job_points, b.job_cap, (const PassCounters*)nullptr, 0xFFFFFFFFu);
    hipLaunchKernelGGL(batch_bin_scatter_kernel, dim3(gx, a.nviews), dim3(256), hist, st, (const BatchView*)a.views,
                       (const uint2*)b.erect, (const uint32_t*)b.ew, b.reg_cursor, b.list, b.lw,
                       (const uint32_t*)b.state);
    const uint32_t grid = b.cus * 4u;
    hipLaunchKernelGGL(bin_tile_sum_kernel, dim3(grid), dim3(256), 0, st, (const uint32_t*)b.lw,
                       (const uint32_t*)b.state, b.tsum);
    hipLaunchKernelGGL(bin_tiles_scan_kernel, dim3(1), dim3(1024), 0, st, (const uint32_t*)b.tsum,
                       (const uint32_t*)b.state, b.toff);
    hipLaunchKernelGGL(bin_local_scan_kernel, dim3(grid), dim3(256), 0, st, (const uint32_t*)b.lw,
                       (const uint32_t*)b.state, (const uint32_t*)b.toff, b.lpre);
    hipLaunchKernelGGL(batch_bin_splat_kernel, dim3(b.cus), dim3(GPS_BIN_THREADS), 0, st, (const SplatRecord*)a.recs,
                       (const uint32_t*)a.cidx, (const uint32_t*)a.coff, (const uint32_t*)b.list,
                       (const uint32_t*)b.lpre, (const uint32_t*)b.reg_ecnt, (const uint32_t*)b.reg_ptot,
                       (const uint32_t*)b.reg_eoff, (const uint32_t*)b.reg_poff, (const uint32_t*)b.reg_jobs,
                       (const uint32_t*)b.job_reg, (const uint32_t*)b.job_p0, b.state, (const BatchView*)a.views,
                       a.nviews, reinterpret_cast<unsigned long long*>(a.fb), b.job_points, a.atomics);
}

void launch_batch_bin_big(const BatchArgs& a, hipStream_t st) {
    const BatchBinArgs& b = a.bin;
    if (!b.any) return;
    hipLaunchKernelGGL(batch_bin_big_kernel, dim3(b.cus * 4u), dim3(256), 0, st, (const SplatRecord*)a.recs,
                       (const uint32_t*)a.cidx, (const uint32_t*)a.coff, (const uint8_t*)a.cview,
                       (const uint32_t*)b.ew, (const uint32_t*)b.big, b.state, (const BatchView*)a.views,
                       reinterpret_cast<unsigned long long*>(a.fb), a.atomics);
}

// ------------------------------------------------------------------ binned splat launcher
void launch_binned_splat(const BinnedArgs& a, hipStream_t st) {
    launch_zero_u32(a.state, 8, st);
    const uint32_t grid = a.cus * 4u;
    const size_t hist = 2u * a.g.nregions * sizeof(uint32_t);
    hipLaunchKernelGGL(bin_count_kernel, dim3(grid), dim3(256), hist, st, a.recs, a.cidx, a.coff, a.ctr, a.g, a.erect,
                       a.ew, a.reg_cnt, a.reg_pts, a.big, a.state, a.dense_points);
    hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, st, a.g, a.reg_cnt, a.reg_pts, a.reg_ecnt, a.reg_ptot,
                       a.reg_eoff, a.reg_poff, a.reg_cursor, a.reg_jobs, a.job_reg, a.job_p0, a.state, a.list_cap,
                       a.job_points, a.job_cap, a.ctr, a.dense_points);
    hipLaunchKernelGGL(bin_scatter_kernel, dim3(grid), dim3(256), hist, st, a.ctr, a.g, (const uint2*)a.erect,
                       (const uint32_t*)a.ew, a.reg_cursor, a.list, a.lw, (const uint32_t*)a.state);
    hipLaunchKernelGGL(bin_tile_sum_kernel, dim3(grid), dim3(256), 0, st, (const uint32_t*)a.lw,
                       (const uint32_t*)a.state, a.tsum);
    hipLaunchKernelGGL(bin_tiles_scan_kernel, dim3(1), dim3(1024), 0, st, (const uint32_t*)a.tsum,
                       (const uint32_t*)a.state, a.toff);
    hipLaunchKernelGGL(bin_local_scan_kernel, dim3(grid), dim3(256), 0, st, (const uint32_t*)a.lw,
                       (const uint32_t*)a.state, (const uint32_t*)a.toff, a.lpre);
    BinSplat b;
    b.g = a.g;
    b.K = a.sc.K;
    b.unbiased = a.sc.unbiased;
    b.frame_seed = a.sc.frame_seed;
    b.background = a.background;
    hipLaunchKernelGGL(bin_splat_kernel, dim3(a.cus), dim3(GPS_BIN_THREADS), 0, st, a.recs, a.cidx, a.coff,
                       (const uint32_t*)a.list, (const uint32_t*)a.lpre, (const uint32_t*)a.reg_ecnt,
                       (const uint32_t*)a.reg_ptot, (const uint32_t*)a.reg_eoff, (const uint32_t*)a.reg_poff,
                       (const uint32_t*)a.reg_jobs, (const uint32_t*)a.job_reg, (const uint32_t*)a.job_p0, a.state,
                       reinterpret_cast<unsigned long long*>(a.fb), b, a.job_points, a.atomics_count);
    hipLaunchKernelGGL(bin_big_kernel, dim3(grid), dim3(256), 0, st, a.recs, a.cidx, a.coff, (const uint32_t*)a.ew,
                       (const uint32_t*)a.big, a.state, reinterpret_cast<unsigned long long*>(a.fb), a.sc,
                       a.atomics_count);
}

uint32_t binned_dense_points(int up_w, int up_h, int K) {
    const uint64_t p = (uint64_t)GPS_BIN_DENSE_SAMPLES * (uint64_t)up_w * (uint64_t)up_h / (uint64_t)(K > 0 ? K : 1);
    return p < 0xFFFFFFFFull ? (uint32_t)p : 0xFFFFFFFFu;
}

bool binned_geometry(int up_w, int up_h, int spp, BinGeom* g) {
    if (spp < 1 || spp > 4) return false;
    g->spp = spp;
    g->up_w = up_w;
    g->up_h = up_h;
    g->rw_shift = spp <= 2 ? 7 : 6;
    g->rh_shift = spp == 1 ? 7 : 6;
    g->rtx = 1 << (g->rw_shift - GPS_TILE_BITS);
    g->rty = 1 << (g->rh_shift - GPS_TILE_BITS);
    g->nrx = (up_w + (1 << g->rw_shift) - 1) >> g->rw_shift;
    g->nry = (up_h + (1 << g->rh_shift) - 1) >> g->rh_shift;
    g->tiles_per_row = fb_tiles_per_row(up_w);
    const uint64_t n = (uint64_t)g->nrx * (uint64_t)g->nry;
    if (n == 0 || n > GPS_BIN_HIST) return false;
    g->nregions = (uint32_t)n;
    return (uint64_t)g->rtx * g->rty * GPS_TILE_WORDS * (uint64_t)spp <= GPS_BIN_WORDS;
}

}  // namespace gps
