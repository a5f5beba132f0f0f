// gps_oracle.cpp — CPU oracle: a plain C++ transcription of the reference's
// per-frame hot path (GaussianPointSplatting::render and what it calls).
//
// TEST INFRASTRUCTURE.  This file is the checker and the reported CPU
// baseline; the product library never links or loads it.
//
// Structure follows the reference one-to-one:
//   render()          src/core/rendering/passes/gaussian_point_splatting.cu:599-737
//   preprocess()      gaussian_point_splatting.cu:147-307
//   build()           src/core/random/workload/workload_distributor.cu:62-145
//                     (masked exclusive scan -> scatter -> inclusive max-scan)
//   splat()           gaussian_point_splatting.cu:383-432 + 67-141
//                     (projection recomputed per point, as the reference does)
//   combine()         gaussian_point_splatting.cu:485-548
//   Hi-Z              src/core/rendering/occlusion/depth_mip_chain.{cu,cuh}
//   block culling     src/core/rendering/occlusion/gaussian_bvh.cu:111-144,171-195,215-329
//   resolve           src/core/rendering/passes/post_processing_kernels.cuh:108-285
// Parallel loops use std::thread over index ranges; the framebuffer min uses a
// CAS loop on std::atomic<uint64_t>, which is order-independent, so results
// equal a serial run.  The biased splat (use_unbiased_2d_splatting = 0) runs
// the reference's pre-check + atomicAddColor per sample on the same atomics;
// its words are order-independent unless orc_pass_biased_events reports a
// zero-colour sample or a colour-sum carry out of bit 35.
//
// Defined behaviours where the reference reads uninitialised memory:
//   * the Hi-Z chain starts at +inf (nothing occluded before the first build);
//   * masks start at 0.
#include "gps_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "orc_ref.h"

using namespace orc;

namespace {

template <typename F> void parallel_for(uint64_t n, int nthreads, uint64_t align, F fn) {
    if (nthreads <= 1 || n < 4096) {
        fn(0, n);
        return;
    }
    uint64_t chunk = (n + nthreads - 1) / nthreads;
    chunk = (chunk + align - 1) / align * align;
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; ++t) {
        const uint64_t b = (uint64_t)t * chunk;
        if (b >= n) break;
        const uint64_t e = std::min(n, b + chunk);
        ts.emplace_back([=]() { fn(b, e); });
    }
    for (auto& t : ts) t.join();
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// IEEE half -> float (exact).
float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1Fu;
    const uint32_t man = h & 0x3FFu;
    float v;
    if (exp == 0) {
        v = (float)man * 0x1.0p-24f;  // subnormal (and zero)
        return sign ? -v : v;
    }
    if (exp == 31) {
        const uint32_t bits = sign | 0x7F800000u | (man << 13);
        return __builtin_bit_cast(float, bits);
    }
    const uint32_t bits = sign | ((exp - 15 + 127) << 23) | (man << 13);
    return __builtin_bit_cast(float, bits);
}
// OCP FP8 E4M3 (bias 7, no infinities, 0x7F/0xFF = NaN) -> float (exact).
float fp8e4m3_to_float(uint8_t q) {
    const bool neg = (q & 0x80u) != 0;
    const uint32_t exp = (q >> 3) & 0xFu;
    const uint32_t man = q & 0x7u;
    float v;
    if (exp == 15 && man == 7) return __builtin_bit_cast(float, 0x7FC00000u);
    if (exp == 0)
        v = (float)man * 0x1.0p-9f;  // 2^-6 * man/8
    else
        v = (1.0f + (float)man * 0.125f) * std::ldexp(1.0f, (int)exp - 7);
    return neg ? -v : v;
}

struct Node {
    vec3<float> mn, mx;
    uint32_t begin, count;
};

struct HiZ {
    int width = 0, height = 0, levels = 0;
    std::vector<int> lw, lh;
    std::vector<size_t> off;
    std::vector<float> mip;

    void init(int w, int h) {
        width = w;
        height = h;
        int maxDim = std::max(w, h), e = 0;
        while ((1 << e) < maxDim) ++e;
        levels = e + 1;
        lw.resize(levels);
        lh.resize(levels);
        off.resize(levels);
        size_t total = 0;
        for (int l = 0; l < levels; ++l) {
            const int fe = levels - 1;
            lw[l] = std::max(1, (w + (1 << (fe - l)) - 1) >> (fe - l));
            lh[l] = std::max(1, (h + (1 << (fe - l)) - 1) >> (fe - l));
            off[l] = total;
            total += (size_t)lw[l] * lh[l];
        }
        mip.assign(total, __builtin_bit_cast(float, 0x7F800000u));
    }
    // depth_mip_chain.cu:7-22,24-67,107-145
    void build(const float* depth, int K_min_levels) {
        const int fl = levels - 1;
        for (int y = 0; y < lh[fl]; ++y)
            for (int x = 0; x < lw[fl]; ++x) mip[off[fl] + (size_t)y * lw[fl] + x] = depth[(size_t)y * lw[fl] + x];
        for (int l = levels - 2; l >= 0; --l) {
            const bool useMin = ((levels - 2) - l) < K_min_levels;
            for (int y = 0; y < lh[l]; ++y)
                for (int x = 0; x < lw[l]; ++x) {
                    float v = useMin ? __builtin_bit_cast(float, 0x7F800000u) : __builtin_bit_cast(float, 0xFF800000u);
                    for (int dy = 0; dy < 2; ++dy)
                        for (int dx = 0; dx < 2; ++dx) {
                            const int fx = 2 * x + dx, fy = 2 * y + dy;
                            if (fx < lw[l + 1] && fy < lh[l + 1]) {
                                const float c = mip[off[l + 1] + (size_t)fy * lw[l + 1] + fx];
                                v = useMin ? gps::det_fmin(v, c) : gps::det_fmax(v, c);
                            }
                        }
                    mip[off[l] + (size_t)y * lw[l] + x] = v;
                }
        }
    }
    // depth_mip_chain.cuh:172-215
    bool is_occluded(float aabbDepth, ivec2_ minPix, ivec2_ maxPix) const {
        minPix.x = std::max(0, minPix.x);
        minPix.y = std::max(0, minPix.y);
        maxPix.x = std::min(width - 1, maxPix.x);
        maxPix.y = std::min(height - 1, maxPix.y);
        const int w = maxPix.x - minPix.x + 1;
        const int h = maxPix.y - minPix.y + 1;
        const int d = std::max(w, h);
        int lod = gps::det_f2i_rz(std::ceil(std::log2((float)d / 2.0f)));
        lod = std::min(std::max(lod, 0), levels - 1);
        const int l = (levels - 1) - lod;
        const int wl = lw[l], hl = lh[l];
        const int tileW = std::max(1, std::min(width, 1 << (levels - l - 1)));
        const int tileH = std::max(1, std::min(height, 1 << (levels - l - 1)));
        const int tx0 = std::max(0, std::min(wl - 1, minPix.x / tileW));
        const int ty0 = std::max(0, std::min(hl - 1, minPix.y / tileH)) * wl;
        const int tx1 = std::max(0, std::min(wl - 1, maxPix.x / tileW));
        const int ty1 = std::max(0, std::min(hl - 1, maxPix.y / tileH)) * wl;
        const size_t b = off[l];
        const float zMax = gps::det_fmax(mip[b + tx0 + ty0],
                                         gps::det_fmax(mip[b + tx1 + ty0], gps::det_fmax(mip[b + tx0 + ty1], mip[b + tx1 + ty1])));
        return zMax < aabbDepth;
    }
};

struct Frame {
    mat4<float> vMatrix, camMatrix, vpMatrix, proj;
    vec2<float> tan_fov, focal;
    vec3<float> campos;
    int up_w, up_h, W, H, ss;
    float near_plane, far_plane;
    const gps_frame_params* p;
};

}  // namespace

struct orc_pass {
    // scene (decoded to fp32)
    uint64_t N = 0;
    std::vector<float> means, rots, scales, opac, sh48;
    // colour from the compressed scene's Sloan-basis evaluation
    // (computeColorFromCompactSH) instead of computeColorFromSHFloat
    bool compact_sh = false;
    template <typename T> vec3<T> sh_color(vec3<T> pos, vec3<T> campos, const T* shs) const {
        return compact_sh ? computeColorFromCompactSH(pos, campos, shs) : computeColorFromSHFloat(pos, campos, shs);
    }
    // settings
    int W = 0, H = 0, spp = 1, ss_created = 1;
    uint32_t maxPoints = 250000000u;
    bool hier_available = true;
    uint32_t gpb = 4;
    int nthreads = 1;
    // workload distributor state
    std::vector<uint32_t> weights, maskA, maskB, indices, masked_weights_dbg;
    uint32_t iteration = 0, num_samples = 0;
    std::vector<vec3<float>> colors;
    // framebuffers
    std::vector<uint64_t> fbA, fbB;
    uint64_t fbSize = 0;
    int fb_ss = -1;
    uint64_t* cur_fb = nullptr;
    // occlusion
    HiZ hiz;
    std::vector<Node> nodes;
    std::vector<uint32_t> culled;
    // stats
    gps_stats stats{};
    double stage_ms[8] = {0};

    std::vector<uint32_t>& maskCurrent() { return iteration % 2 == 0 ? maskA : maskB; }
    std::vector<uint32_t>& maskPrevious() { return iteration % 2 == 0 ? maskB : maskA; }

    vec3<float> mean(uint64_t g) const { return {means[3 * g], means[3 * g + 1], means[3 * g + 2]}; }
    vec4<float> quat(uint64_t g) const { return {rots[4 * g], rots[4 * g + 1], rots[4 * g + 2], rots[4 * g + 3]}; }
    vec3<float> scale(uint64_t g) const { return {scales[3 * g], scales[3 * g + 1], scales[3 * g + 2]}; }
    // get_opacity: glm::min(opacity, MAX_GAUSSIAN_OPACITY) (gaussian_types.cuh:616-618)
    float opacity(uint64_t g) const { return gmin(opac[g], 1.0f); }

    bool is_culled(uint32_t g) const {
        const uint32_t span = gpb * 32u;
        const uint32_t treeId = g / span, within = g % span, blockId = within / gpb;
        return ((culled[treeId] >> blockId) & 1u) != 0u;
    }
    static bool is_valid(uint32_t idx, const std::vector<uint32_t>& m) { return ((m[idx / 32] >> (idx % 32)) & 1u) != 0; }

    // gaussian_point_splatting.cu:17-65
    bool loadAndTransform(uint32_t g, const Frame& f, vec4<float>& proj_mean, vec2<float>& pixel_mean,
                          vec3<float>& cov2d, ivec2_& pmin, ivec2_& pmax, float& det, float& view_z) const {
        const vec3<float> m = mean(g);
        const vec3<float> view_mean = transformPoint4x3(m, f.vMatrix);
        view_z = view_mean.z;
        if (-view_z <= 0.2f) return false;
        proj_mean = f.vpMatrix * vec4<float>{m.x, m.y, m.z, 1.0f};
        proj_mean = proj_mean * (1.0f / proj_mean.w);
        if (proj_mean.z < -1.0f || proj_mean.z > 1.0f) return false;
        const mat3<float> cov = scale_rot_to_cov(scale(g), quat(g));
        cov2d = computeCov2D(f.vMatrix, cov, m, f.tan_fov, f.focal);
        const float dil = 0.3f * (float)f.ss * (float)f.ss;
        cov2d = cov2d + vec3<float>{dil, 0.0f, dil};
        det = cov2d.x * cov2d.z - cov2d.y * cov2d.y;
        if (det <= 0.0f) return false;
        pixel_mean = (vec2<float>{proj_mean.x, proj_mean.y} * 0.5f + vec2<float>{0.5f, 0.5f}) *
                     vec2<float>{(float)f.up_w, (float)f.up_h};
        if (!compute2DGaussianBounds3DGS(pmin, pmax, f.up_w, f.up_h, cov2d, pixel_mean)) return false;
        return true;
    }

    void make_frame(const gps_frame_params* p, Frame& f) const {
        f.p = p;
        f.vMatrix = mat4_from<float>(p->view);
        f.proj = mat4_from<float>(p->proj);
        f.camMatrix = affine_inverse(f.vMatrix);
        f.vpMatrix = f.proj * f.vMatrix;
        f.W = p->width;
        f.H = p->height;
        f.ss = p->supersampling_factor;
        f.up_w = p->width * p->supersampling_factor;
        f.up_h = p->height * p->supersampling_factor;
        f.near_plane = p->near_plane;
        f.far_plane = p->far_plane;
        const float r = float(p->width) / p->height;
        const float t_f_y = std::tan(p->fovy_degrees * 0.01745329251994329576923690768489f * 0.5f);
        const float fovx = 2.0f * std::atan(t_f_y * r);
        const float t_f_x = std::tan(fovx * 0.5f);
        f.tan_fov = {t_f_x, t_f_y};
        f.focal = {(float)(p->width * p->supersampling_factor) / (2.0f * t_f_x),
                   (float)(p->height * p->supersampling_factor) / (2.0f * t_f_y)};
        f.campos = {f.camMatrix[3][0], f.camMatrix[3][1], f.camMatrix[3][2]};
    }

    void update_image_buffers(int ss) {
        const int uw = W * ss, uh = H * ss;
        const uint64_t tr = (uint64_t)((uw + 63) >> 6), tc = (uint64_t)((uh + 63) >> 6);
        const uint64_t want = tr * tc * 4096ull * (uint64_t)spp;
        if (want == fbSize) return;
        fbSize = want;
        fbA.assign(want, 0);
        fbB.assign(want, 0);
    }

    // gaussian_bvh.cu:287-329
    void compute_occlusion_mask(const Frame& f, bool occl) {
        const uint32_t words = (uint32_t)culled.size();
        const float* pl = f.p->frustum_planes;
        parallel_for(words, nthreads, 1, [&](uint64_t b, uint64_t e) {
            for (uint64_t w = b; w < e; ++w) {
                uint32_t word = 0;
                for (uint32_t lane = 0; lane < 32; ++lane) {
                    const uint32_t idx = (uint32_t)w * 32 + lane;
                    bool c = false;
                    if (idx < nodes.size()) {
                        const Node& n = nodes[idx];
                        bool outside = false;
                        for (int i = 0; i < 6 && !outside; ++i) {
                            const vec3<float> nn = {pl[4 * i], pl[4 * i + 1], pl[4 * i + 2]};
                            const float d = pl[4 * i + 3];
                            const vec3<float> v = {nn.x >= 0.0f ? n.mx.x : n.mn.x, nn.y >= 0.0f ? n.mx.y : n.mn.y,
                                                   nn.z >= 0.0f ? n.mx.z : n.mn.z};
                            if (dot(nn, v) + d < 0.0f) outside = true;
                        }
                        if (outside)
                            c = true;
                        else if (occl) {
                            ivec2_ mnp, mxp;
                            float fd;
                            if (project_bounds(n, f.vpMatrix, hiz.width, hiz.height, mnp, mxp, fd))
                                c = hiz.is_occluded(fd, mnp, mxp);
                        }
                    }
                    if (c) word |= 1u << lane;
                }
                culled[w] = word;
            }
        });
    }
    // gaussian_bvh.cu:237-285
    static bool project_bounds(const Node& n, const mat4<float>& vp, int width, int height, ivec2_& outMin,
                               ivec2_& outMax, float& outFront) {
        const vec3<float> b0 = n.mn, b1 = n.mx;
        const vec3<float> corners[8] = {{b0.x, b0.y, b0.z}, {b1.x, b0.y, b0.z}, {b0.x, b1.y, b0.z}, {b1.x, b1.y, b0.z},
                                        {b0.x, b0.y, b1.z}, {b1.x, b0.y, b1.z}, {b0.x, b1.y, b1.z}, {b1.x, b1.y, b1.z}};
        int minx = width - 1, miny = height - 1, maxx = 0, maxy = 0;
        float minDepth = 1e30f;
        bool anyFront = false;
        for (int i = 0; i < 8; ++i) {
            const vec4<float> clip = vp * vec4<float>{corners[i].x, corners[i].y, corners[i].z, 1.0f};
            if (clip.w <= 0.0f) continue;
            anyFront = true;
            const vec3<float> ndc = vec3<float>{clip.x, clip.y, clip.z} / clip.w;
            const int px = gps::det_f2i_rz((ndc.x * 0.5f + 0.5f) * width);
            const int py = gps::det_f2i_rz((ndc.y * 0.5f + 0.5f) * height);
            minx = std::min(minx, px);
            miny = std::min(miny, py);
            maxx = std::max(maxx, px);
            maxy = std::max(maxy, py);
            minDepth = gps::det_fmin(minDepth, ndc.z);
        }
        if (!anyFront) return false;
        minx = std::max(minx, 0);
        miny = std::max(miny, 0);
        maxx = std::min(maxx, width - 1);
        maxy = std::min(maxy, height - 1);
        if (minx > maxx || miny > maxy) return false;
        outMin = {minx, miny};
        outMax = {maxx, maxy};
        outFront = minDepth;
        return true;
    }

    // gaussian_point_splatting.cu:147-307
    void preprocess(const Frame& f, bool onlyPrev, int frame_seed, bool hier) {
        const gps_frame_params* p = f.p;
        const bool occl = p->enable_occlusion_culling != 0;
        const bool unbiased = p->use_unbiased_2d_splatting != 0;
        std::vector<uint32_t>& mcur = maskCurrent();
        const std::vector<uint32_t>& mprev = maskPrevious();
        const uint64_t nwarps = (N + 31) / 32;
        const int slot = onlyPrev ? 1 : 0;
        for (auto& r : rejects[slot]) r = 0;
        parallel_for(nwarps, nthreads, 1, [&](uint64_t wb, uint64_t we) {
            // why each Gaussian of an unculled block is not valid (counted
            // per range, then added): tested, valid in the first pass,
            // opacity, projection, off screen, small, Hi-Z, count 0, valid
            uint64_t rj[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (uint64_t w = wb; w < we; ++w) {
                uint32_t ballot = 0;
                for (uint32_t lane = 0; lane < 32; ++lane) {
                    const uint64_t idx64 = w * 32 + lane;
                    if (idx64 >= N) break;
                    const uint32_t idx = (uint32_t)idx64;
                    uint2_ seed = makeSeed(idx, (uint32_t)frame_seed);
                    bool splats = false;
                    int reason = 0;
                    do {
                        if (hier && is_culled(idx)) break;
                        rj[0]++;
                        if (onlyPrev && is_valid(idx, mprev)) { reason = 1; break; }
                        const float op = opacity(idx);
                        if (op <= 0.0f) { reason = 2; break; }
                        vec4<float> proj_mean;
                        vec2<float> pm;
                        vec3<float> cov2d;
                        ivec2_ pmin, pmax;
                        float det, view_z;
                        if (!loadAndTransform(idx, f, proj_mean, pm, cov2d, pmin, pmax, det, view_z)) { reason = 3; break; }
                        if (pmin.x >= f.up_w || pmin.y >= f.up_h || pmax.x < 0 || pmax.y < 0) { reason = 4; break; }
                        if (p->cull_small_gaussians) {
                            const float tr = cov2d.x + cov2d.z;
                            const float t2 = 0.5f * tr;
                            const float term2 = std::sqrt(gps::det_fmax(0.1f, t2 * t2 - det));
                            const float e1 = t2 + term2, e2 = t2 - term2;
                            if (e1 < 0.02f || e2 < 0.02f) { reason = 5; break; }
                        }
                        const float TWO_PI = 2.0f * 3.14159265358979323846f;
                        const float s_det = std::sqrt(det);
                        const float importance = unbiased ? TWO_PI * s_det * dilog(op) : op * TWO_PI * s_det;
                        if (occl && hiz.is_occluded(proj_mean.z, {pmin.x / f.ss, pmin.y / f.ss},
                                                    {pmax.x / f.ss, pmax.y / f.ss})) { reason = 6; break; }
                        float K = (float)(f.ss * f.ss * 1);
                        if (!p->reduce_point_count) K = 1.0f;
                        uint32_t n = unbiased ? poisson(seed, importance)
                                              : (uint32_t)stochastic_round(importance, pcg2d(seed).x);
                        if (K > 1) n = (uint32_t)stochastic_round(static_cast<float>(n) / K, pcg2d(seed).x);
                        if (n <= 0) { reason = 7; break; }
                        n = std::min(n, gps::det_f2u_rz((float)(f.up_w * f.up_h) / (2 * K)));
                        colors[idx] = sh_color(mean(idx), f.campos, &sh48[48 * (uint64_t)idx]);
                        weights[idx] = n;
                        splats = true;
                        reason = 8;
                    } while (false);
                    if (reason) rj[reason]++;
                    if (splats) ballot |= 1u << lane;
                }
                mcur[w] = ballot;
            }
            for (int k = 0; k < 9; ++k) rejects[slot][k] += rj[k];
        });
    }
    std::atomic<uint64_t> rejects[2][9];

    // workload_distributor.cu:62-145
    void build() {
        if (N == 0) {
            num_samples = 0;
            return;
        }
        std::vector<uint32_t>& mcur = maskCurrent();
        uint32_t last_count = weights[N - 1];
        if (!is_valid((uint32_t)(N - 1), mcur)) last_count = 0;
        masked_weights_dbg.assign(N, 0);
        for (uint64_t i = 0; i < N; ++i) masked_weights_dbg[i] = is_valid((uint32_t)i, mcur) ? weights[i] : 0u;
        uint32_t acc = 0;
        for (uint64_t i = 0; i < N; ++i) {
            const uint32_t v = masked_weights_dbg[i];
            weights[i] = acc;
            acc += v;
        }
        const uint32_t last_offset = weights[N - 1];
        num_samples = std::min(last_offset + last_count, maxPoints);
        stats_total_weight = (uint64_t)(uint32_t)(last_offset + last_count);
        if (num_samples <= 0) {
            num_samples = 0;
            return;
        }
        if (indices.size() < num_samples) indices.resize(num_samples);
        std::fill(indices.begin(), indices.begin() + num_samples, 0u);
        for (uint64_t i = 0; i < N; ++i)
            if (is_valid((uint32_t)i, mcur)) {
                const uint32_t pidx = weights[i];
                if (pidx < num_samples) indices[pidx] = (uint32_t)i;
            }
        uint32_t mx = 0;
        for (uint32_t i = 0; i < num_samples; ++i) {
            mx = std::max(mx, indices[i]);
            indices[i] = mx;
        }
        iteration++;
    }
    uint64_t stats_total_weight = 0;
    BiasedEvents biased_events;

    // gaussian_point_splatting.cu:67-141, 383-432
    void splat(const Frame& f, int frame_seed, uint32_t numPoints) {
        const gps_frame_params* p = f.p;
        const bool unbiased = p->use_unbiased_2d_splatting != 0;
        int K = f.ss * f.ss * 1;
        if (!p->reduce_point_count) K = 1;
        uint64_t* fb = cur_fb;
        parallel_for(numPoints, nthreads, 1, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; ++i) {
                const uint32_t idx = (uint32_t)i;
                uint2_ seed = makeSeed(idx, (uint32_t)frame_seed);
                const uint32_t g = indices[idx];
                vec4<float> proj_mean;
                vec2<float> pm;
                vec3<float> cov2d;
                ivec2_ pmin, pmax;
                float det, view_z;
                if (!loadAndTransform(g, f, proj_mean, pm, cov2d, pmin, pmax, det, view_z)) continue;
                const uint64_t color = get_color_packed(colors[g]);
                const float op = opacity(g);
                const uint64_t vdb = convert_view_depth_bits(view_z, f.near_plane, f.far_plane) << 36;
                const vec3<float> con = conic(cov2d, 1.0f / det);
                const float c0 = std::sqrt(cov2d.x);
                const float c1 = cov2d.y / c0;
                const float c2 = std::sqrt(cov2d.z - c1 * c1);
                for (int q = 0; q < spp * K; ++q) {
                    const float2_ u = pcg2d(seed);
                    const float2_ xy = unbiased ? correctedBoxMuller(u.x, u.y, op) : boxMuller(u.x, u.y);
                    vec2<float> pixel = {pm.x + c0 * xy.x, pm.y + c1 * xy.x + c2 * xy.y};
                    pixel = {std::floor(pixel.x), std::floor(pixel.y)};
                    const int x = gps::det_f2i_rz(pixel.x);
                    const int y = gps::det_f2i_rz(pixel.y);
                    const vec2<float> delta = pm - pixel;
                    if (x < pmin.x || x >= pmax.x || y < pmin.y || y >= pmax.y) continue;
                    const float gw = gaussian_value(con, delta);
                    const float alpha = op * gw;
                    if (alpha < 1.0f / 255.0f) continue;
                    const int index = posToMemory(x, y, q / K, spp, f.up_w, f.up_h);
                    const uint64_t val = vdb | color;
                    std::atomic_ref<uint64_t> ref(fb[index]);
                    if (!unbiased) {
                        // :130-138, atomicAddColor branch
                        splat_sample_biased(ref, val, &biased_events);
                        continue;
                    }
                    // :130-135: the pre-check, then atomicMin (CAS loop)
                    uint64_t old = ref.load(std::memory_order_relaxed);
                    if (old <= vdb) continue;
                    while (val < old && !ref.compare_exchange_weak(old, val, std::memory_order_relaxed)) {
                    }
                }
            }
        });
    }

    // gaussian_point_splatting.cu:485-548
    void combine(bool depthOnly, const Frame& f, float* color, float* depth, uint64_t initial) {
        const int ssf = f.ss;
        parallel_for((uint64_t)f.H, nthreads, 1, [&](uint64_t yb, uint64_t ye) {
            for (int y = (int)yb; y < (int)ye; ++y)
                for (int x = 0; x < f.W; ++x) {
                    const uint32_t idx = (uint32_t)(y * f.W + x);
                    const int bx = x * ssf, by = y * ssf;
                    float d = 1e9f;
                    vec3<float> c = {0.0f, 0.0f, 0.0f};
                    for (int dx = 0; dx < ssf; dx++)
                        for (int dy = 0; dy < ssf; dy++)
                            for (int pp = 0; pp < spp; ++pp) {
                                const int i = posToMemory(bx + dx, by + dy, pp, spp, f.up_w, f.up_h);
                                if ((uint64_t)i >= fbSize) continue;
                                const uint64_t val = cur_fb[i];
                                (void)initial;
                                const float vd = convert_view_depth_float((val >> 36ull) & ((1ull << 28) - 1),
                                                                          f.near_plane, f.far_plane);
                                const float z = view_to_projected_depth(vd, f.near_plane, f.far_plane);
                                d = gmin(d, z);
                                if (depthOnly) continue;
                                c.x += static_cast<float>((val >> 24) & 0xFFF) / 255.0f;
                                c.y += static_cast<float>((val >> 12) & 0xFFF) / 255.0f;
                                c.z += static_cast<float>((val >> 0) & 0xFFF) / 255.0f;
                            }
                    const float inv = 1.0f / (spp * ssf * ssf);
                    if (!depthOnly) {
                        color[3 * idx + 0] = c.x * inv;
                        color[3 * idx + 1] = c.y * inv;
                        color[3 * idx + 2] = c.z * inv;
                    }
                    depth[idx] = d;
                }
        });
    }

    int render(const gps_frame_params* p, float* color, float* depth) {
        Frame f;
        make_frame(p, f);
        if (p->width != W || p->height != H) return GPS_ERR_INVALID_ARG;
        update_image_buffers(p->supersampling_factor);
        const uint64_t bg = 0xFFFFFFF000000000ull | vec3To36Bits({p->background[0], p->background[1], p->background[2]});
        cur_fb = (p->frame % 2 == 0) ? fbA.data() : fbB.data();
        for (double& s : stage_ms) s = 0;
        double t0 = now_ms();
        std::fill(cur_fb, cur_fb + fbSize, bg);
        stage_ms[0] += now_ms() - t0;
        stats = gps_stats{};
        stats.num_gaussians = N;
        biased_events.zero_colour = 0;
        biased_events.wrapped = 0;
        for (auto& ps : rejects)
            for (auto& r : ps) r = 0;
        int K = f.ss * f.ss;
        if (!p->reduce_point_count) K = 1;
        stats.K = K;
        auto renderPass = [&](bool first) {
            const int fs = 4 * p->frame + (first ? 0 : 2);
            const bool hier = p->enable_hierarchical_culling && hier_available;
            double t = now_ms();
            if (hier) compute_occlusion_mask(f, p->enable_occlusion_culling != 0);
            stage_ms[1] += now_ms() - t;
            t = now_ms();
            preprocess(f, !first, fs, hier);
            stage_ms[2] += now_ms() - t;
            t = now_ms();
            build();
            stage_ms[3] += now_ms() - t;
            const int slot = first ? 0 : 1;
            stats.num_points[slot] = num_samples;
            stats.total_weight[slot] = stats_total_weight;
            if (stats_total_weight > num_samples) stats.budget_truncated = 1;
            uint64_t v = 0;
            const std::vector<uint32_t>& m = iteration_mask_of_last_pass();
            for (uint64_t i = 0; i < m.size(); ++i) v += (uint64_t)__builtin_popcount(m[i]);
            stats.valid_gaussians[slot] = v;
            if (hier) {
                uint64_t cb = 0;
                for (uint32_t wd : culled) cb += (uint64_t)__builtin_popcount(wd);
                stats.culled_blocks[slot] = cb;
            }
            stats.samples += (uint64_t)num_samples * (uint64_t)K * (uint64_t)spp;
            stats.passes = first ? 1 : 2;
            t = now_ms();
            if (num_samples > 0) splat(f, fs + 1, num_samples);
            stage_ms[4] += now_ms() - t;
        };
        renderPass(true);
        if (p->enable_occlusion_culling) {
            double t = now_ms();
            combine(true, f, color, depth, bg);
            stage_ms[5] += now_ms() - t;
            t = now_ms();
            hiz.build(depth, p->occlusion_culling_min_reduce_count);
            stage_ms[6] += now_ms() - t;
            renderPass(false);
        }
        double t = now_ms();
        combine(false, f, color, depth, bg);
        stage_ms[5] += now_ms() - t;
        if (p->enable_occlusion_culling) {
            t = now_ms();
            hiz.build(depth, p->occlusion_culling_min_reduce_count);
            stage_ms[6] += now_ms() - t;
        }
        return GPS_OK;
    }
    // Mask written by the most recent preprocess (build() flips the iteration
    // only when points were produced).
    const std::vector<uint32_t>& iteration_mask_of_last_pass() {
        return (num_samples > 0) ? maskPrevious() : maskCurrent();
    }

    void init_common(const gps_render_settings* s, int nt) {
        W = s->width;
        H = s->height;
        spp = s->render_pass_count;
        ss_created = s->supersampling_factor;
        maxPoints = (uint32_t)std::min<int64_t>(s->max_number_of_points, 0xFFFFFFFFll);
        hier_available = s->fully_disable_hierarchical_culling == 0;
        nthreads = std::max(1, nt);
        weights.assign(N, 0);
        maskA.assign((N + 31) / 32, 0);
        maskB.assign((N + 31) / 32, 0);
        colors.assign(N, vec3<float>{0, 0, 0});
        hiz.init(W, H);
        update_image_buffers(ss_created);
        // gaussian_bvh.cu:171-195 (blocks over the current order), :111-144
        if (s->gaussians_per_block > 0)
            gpb = (uint32_t)s->gaussians_per_block;
        else
            gpb = (uint32_t)gmax(4.0f, std::floor((float)N / (float)1e6));
        const uint64_t nb = (N + gpb - 1) / gpb;
        nodes.resize(nb);
        parallel_for(nb, nthreads, 1, [&](uint64_t b0, uint64_t b1) {
            for (uint64_t b = b0; b < b1; ++b) {
                vec3<float> bmin = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
                vec3<float> bmax = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
                const uint64_t e = std::min<uint64_t>(N, (b + 1) * gpb);
                for (uint64_t i = b * gpb; i < e; ++i) {
                    const mat3<float> cov = scale_rot_to_cov(scale(i), quat(i));
                    const vec3<float> ext =
                        3.0f * vec3<float>{std::sqrt(cov[0].x), std::sqrt(cov[1].y), std::sqrt(cov[2].z)};
                    bmin = vmin(bmin, mean(i) - ext);
                    bmax = vmax(bmax, mean(i) + ext);
                }
                nodes[b] = {bmin, bmax, (uint32_t)(b * gpb), (uint32_t)(e - b * gpb)};
            }
        });
        culled.assign((nb + 31) / 32, 0);
    }
};

extern "C" {

orc_pass* orc_pass_create(const gps_render_settings* s, const gps_scene_desc* d, int32_t nthreads) {
    if (!s || !d || d->format != GPS_FORMAT_FP32) return nullptr;
    orc_pass* p = new orc_pass();
    const uint64_t N = d->num_gaussians;
    p->N = N;
    p->means.assign(d->means, d->means + 3 * N);
    p->rots.assign(d->rotations, d->rotations + 4 * N);
    p->scales.assign(d->scales, d->scales + 3 * N);
    p->opac.assign(d->opacities, d->opacities + N);
    p->sh48.assign(48 * N, 0.0f);
    const int nc = (d->sh_degree + 1) * (d->sh_degree + 1);
    for (uint64_t g = 0; g < N; ++g)
        for (int k = 0; k < 3 * nc; ++k) p->sh48[48 * g + k] = d->shs[(uint64_t)3 * nc * g + k];
    p->init_common(s, nthreads);
    return p;
}

orc_pass* orc_pass_create_half_fp8(const gps_render_settings* s, const orc_half_fp8_scene* d, int32_t nthreads) {
    if (!s || !d) return nullptr;
    orc_pass* p = new orc_pass();
    const uint64_t N = d->num_gaussians;
    p->N = N;
    p->means.assign(d->means, d->means + 3 * N);
    p->rots.resize(4 * N);
    p->scales.resize(3 * N);
    p->opac.resize(N);
    p->sh48.assign(48 * N, 0.0f);
    const int nrest = (d->sh_degree + 1) * (d->sh_degree + 1) - 1;
    parallel_for(N, std::max(1, nthreads), 1, [&](uint64_t g0, uint64_t g1) {
        for (uint64_t g = g0; g < g1; ++g) {
            for (int k = 0; k < 4; ++k) p->rots[4 * g + k] = half_to_float(d->rotations[4 * g + k]);
            for (int k = 0; k < 3; ++k) p->scales[3 * g + k] = half_to_float(d->scales[3 * g + k]);
            p->opac[g] = half_to_float(d->opacities[g]);
            for (int k = 0; k < 3; ++k) p->sh48[48 * g + k] = half_to_float(d->sh_dc[3 * g + k]);
            for (int k = 0; k < 3 * nrest; ++k)
                p->sh48[48 * g + 3 + k] = fp8e4m3_to_float(d->sh_rest[45 * g + k]) * d->sh_scale[g];
        }
    });
    p->init_common(s, nthreads);
    return p;
}

// GPUGaussianSceneCompressed (gaussian_types.cuh:421-541): every Gaussian
// decoded once (decode_gaussian_geometry, get_opacity, CompactSH::get); the
// colour is then the Sloan-basis evaluation.
orc_pass* orc_pass_create_quantized(const gps_render_settings* s, const orc_quantized_scene* d, int32_t nthreads) {
    if (!s || !d) return nullptr;
    orc_pass* p = new orc_pass();
    const uint64_t N = d->num_gaussians;
    p->N = N;
    p->means.resize(3 * N);
    p->rots.resize(4 * N);
    p->scales.resize(3 * N);
    p->opac.resize(N);
    p->sh48.assign(48 * N, 0.0f);
    for (uint64_t g = 0; g < N; ++g) {
        float out[59];
        orc_quantized_decode(d, g, out);
        std::copy(out, out + 3, &p->means[3 * g]);
        std::copy(out + 3, out + 7, &p->rots[4 * g]);
        std::copy(out + 7, out + 10, &p->scales[3 * g]);
        p->opac[g] = out[10];
        std::copy(out + 11, out + 59, &p->sh48[48 * g]);
    }
    p->compact_sh = true;
    p->init_common(s, nthreads);
    return p;
}

void orc_quantized_decode(const orc_quantized_scene* d, uint64_t g, float* out) {
    const GaussianGeometry geo = decode_gaussian_geometry(d->geometry + 4 * g, d->bounds);
    out[0] = geo.mean.x;
    out[1] = geo.mean.y;
    out[2] = geo.mean.z;
    out[3] = geo.quat.x;
    out[4] = geo.quat.y;
    out[5] = geo.quat.z;
    out[6] = geo.quat.w;
    out[7] = geo.scale.x;
    out[8] = geo.scale.y;
    out[9] = geo.scale.z;
    out[10] = normalized_byte_to_float(d->opacity[g]);
    for (int i = 0; i < 16; ++i) {
        const vec3<float> c = compact_sh_get(i, d->sh_dc + 3 * g, d->sh_codes + 45 * g, d->bounds);
        out[11 + 3 * i] = c.x;
        out[12 + 3 * i] = c.y;
        out[13 + 3 * i] = c.z;
    }
}

void orc_pass_destroy(orc_pass* p) { delete p; }
void orc_pass_set_threads(orc_pass* p, int32_t n) { p->nthreads = std::max(1, (int)n); }

int32_t orc_pass_render(orc_pass* p, const gps_frame_params* f, float* color, float* depth) {
    if (!p || !f || !color || !depth) return GPS_ERR_INVALID_ARG;
    return p->render(f, color, depth);
}

int32_t orc_pass_stats(orc_pass* p, gps_stats* out) {
    *out = p->stats;
    return GPS_OK;
}

void orc_pass_biased_events(orc_pass* p, uint64_t out[2]) {
    out[0] = p->biased_events.zero_colour.load();
    out[1] = p->biased_events.wrapped.load();
}

void orc_pass_rejects(orc_pass* p, uint64_t out[18]) {
    for (int s = 0; s < 2; ++s)
        for (int k = 0; k < 9; ++k) out[9 * s + k] = p->rejects[s][k].load();
}

uint64_t orc_biased_sample(uint64_t* word, uint64_t value, int32_t precheck, int32_t add_if_equal) {
    std::atomic_ref<uint64_t> ref(*word);
    if (precheck) {
        splat_sample_biased(ref, value);
        return 1;
    }
    return atomicAddColor(ref, value, add_if_equal != 0) ? 1 : 0;
}

void orc_pass_stage_ms(orc_pass* p, double out[8]) {
    for (int i = 0; i < 8; ++i) out[i] = p->stage_ms[i];
}

uint64_t orc_pass_debug_size(orc_pass* p, int32_t which) {
    switch (which) {
        case GPS_DBG_WEIGHTS: return p->N;
        case GPS_DBG_MASK_CURRENT:
        case GPS_DBG_MASK_PREVIOUS: return (p->N + 31) / 32;
        case GPS_DBG_INDICES: return p->num_samples;
        case GPS_DBG_FRAMEBUFFER: return p->fbSize;
        case GPS_DBG_RECORDS: return 16 * p->N;
        case GPS_DBG_HIZ: return p->hiz.mip.size();
        case GPS_DBG_CULL_MASK: return p->culled.size();
        case GPS_DBG_VALID_IDS:
        case GPS_DBG_VALID_OFFSETS: {
            uint64_t v = 0;
            for (uint64_t i = 0; i < p->N; ++i) v += p->masked_weights_dbg.size() == p->N && p->masked_weights_dbg[i] > 0;
            return v;
        }
    }
    return 0;
}

int32_t orc_pass_debug_copy(orc_pass* p, int32_t which, void* dst, uint64_t count) {
    if (count > orc_pass_debug_size(p, which)) return GPS_ERR_INVALID_ARG;
    switch (which) {
        case GPS_DBG_WEIGHTS:
            std::memcpy(dst, p->masked_weights_dbg.data(), count * 4);
            return GPS_OK;
        case GPS_DBG_MASK_CURRENT:
            std::memcpy(dst, p->iteration_mask_of_last_pass().data(), count * 4);
            return GPS_OK;
        case GPS_DBG_MASK_PREVIOUS: {
            const auto& cur = p->iteration_mask_of_last_pass();
            const auto& other = (&cur == &p->maskA) ? p->maskB : p->maskA;
            std::memcpy(dst, other.data(), count * 4);
            return GPS_OK;
        }
        case GPS_DBG_INDICES:
            std::memcpy(dst, p->indices.data(), count * 4);
            return GPS_OK;
        case GPS_DBG_FRAMEBUFFER:
            std::memcpy(dst, p->cur_fb, count * 8);
            return GPS_OK;
        case GPS_DBG_HIZ:
            std::memcpy(dst, p->hiz.mip.data(), count * 4);
            return GPS_OK;
        case GPS_DBG_CULL_MASK:
            std::memcpy(dst, p->culled.data(), count * 4);
            return GPS_OK;
        case GPS_DBG_VALID_IDS:
        case GPS_DBG_VALID_OFFSETS: {
            uint32_t* o = (uint32_t*)dst;
            uint64_t k = 0;
            for (uint64_t i = 0; i < p->N && k < count; ++i)
                if (p->masked_weights_dbg[i] > 0) o[k++] = (which == GPS_DBG_VALID_IDS) ? (uint32_t)i : p->weights[i];
            return GPS_OK;
        }
        case GPS_DBG_RECORDS:
            return GPS_ERR_UNSUPPORTED;
    }
    return GPS_ERR_INVALID_ARG;
}

int32_t orc_pass_set_state(orc_pass* p, const float* hiz, uint64_t hiz_count, const uint32_t* mask_other,
                           uint64_t mask_words) {
    if (!p || !hiz || !mask_other) return GPS_ERR_INVALID_ARG;
    if (hiz_count != p->hiz.mip.size() || mask_words != p->maskPrevious().size()) return GPS_ERR_INVALID_ARG;
    std::memcpy(p->hiz.mip.data(), hiz, hiz_count * 4);
    std::memcpy(p->maskPrevious().data(), mask_other, mask_words * 4);
    return GPS_OK;
}

int32_t orc_pass_cull(orc_pass* p, const gps_frame_params* fp) {
    if (!p || !fp || p->nodes.empty()) return GPS_ERR_INVALID_ARG;
    Frame f;
    p->make_frame(fp, f);
    p->compute_occlusion_mask(f, fp->enable_occlusion_culling != 0);
    return GPS_OK;
}

int32_t orc_project_ids(orc_pass* p, const gps_frame_params* fp, const uint32_t* ids, uint64_t n, double* out) {
    Frame f;
    p->make_frame(fp, f);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t g = ids[i];
        double* o = out + 16 * i;
        for (int k = 0; k < 16; ++k) o[k] = 0.0;
        if (g >= p->N) continue;
        vec4<float> pmn;
        vec2<float> pm;
        vec3<float> cov;
        ivec2_ a, bb;
        float det, vz;
        if (!p->loadAndTransform(g, f, pmn, pm, cov, a, bb, det, vz)) continue;
        const vec3<float> c = p->sh_color(p->mean(g), f.campos, &p->sh48[48 * (uint64_t)g]);
        const float op = p->opacity(g);
        const float imp = 2.0f * 3.14159265358979323846f * std::sqrt(det) * dilog(op);
        const double vals[16] = {1, vz, pm.x, pm.y, cov.x, cov.y, cov.z, c.x, c.y, c.z, imp, pmn.z,
                                 (double)a.x, (double)a.y, (double)bb.x, (double)bb.y};
        for (int k = 0; k < 16; ++k) o[k] = vals[k];
    }
    return GPS_OK;
}

int32_t orc_project(orc_pass* p, const gps_frame_params* fp, int32_t use_fp64, double* out) {
    Frame f;
    p->make_frame(fp, f);
    const uint64_t N = p->N;
    parallel_for(N, p->nthreads, 1, [&](uint64_t b, uint64_t e) {
        for (uint64_t g = b; g < e; ++g) {
            double* o = out + 16 * g;
            for (int k = 0; k < 16; ++k) o[k] = 0.0;
            if (!use_fp64) {
                vec4<float> pmn;
                vec2<float> pm;
                vec3<float> cov;
                ivec2_ a, bb;
                float det, vz;
                if (!p->loadAndTransform((uint32_t)g, f, pmn, pm, cov, a, bb, det, vz)) continue;
                const vec3<float> c = p->sh_color(p->mean(g), f.campos, &p->sh48[48 * g]);
                const float op = p->opacity(g);
                const float imp = 2.0f * 3.14159265358979323846f * std::sqrt(det) * dilog(op);
                const double vals[16] = {1, vz, pm.x, pm.y, cov.x, cov.y, cov.z, c.x, c.y, c.z, imp, pmn.z,
                                         (double)a.x, (double)a.y, (double)bb.x, (double)bb.y};
                for (int k = 0; k < 16; ++k) o[k] = vals[k];
            } else {
                const mat4<double> V = mat4_from<double>(fp->view), P = mat4_from<double>(fp->proj);
                const mat4<double> VP = P * V;
                const mat4<double> C = affine_inverse(V);
                const vec3<double> campos = {C[3][0], C[3][1], C[3][2]};
                const vec3<double> m = {p->means[3 * g], p->means[3 * g + 1], p->means[3 * g + 2]};
                const vec4<double> q = {p->rots[4 * g], p->rots[4 * g + 1], p->rots[4 * g + 2], p->rots[4 * g + 3]};
                const vec3<double> s = {p->scales[3 * g], p->scales[3 * g + 1], p->scales[3 * g + 2]};
                const vec3<double> vm = transformPoint4x3(m, V);
                if (-vm.z <= 0.2) continue;
                vec4<double> pr = VP * vec4<double>{m.x, m.y, m.z, 1.0};
                pr = pr * (1.0 / pr.w);
                if (pr.z < -1.0 || pr.z > 1.0) continue;
                const double r = double(fp->width) / fp->height;
                const double tfy = std::tan(fp->fovy_degrees * M_PI / 180.0 * 0.5);
                const double tfx = std::tan(std::atan(tfy * r));
                const vec2<double> tf = {tfx, tfy};
                const vec2<double> fo = {fp->width * fp->supersampling_factor / (2.0 * tfx),
                                         fp->height * fp->supersampling_factor / (2.0 * tfy)};
                vec3<double> cov = computeCov2D(V, scale_rot_to_cov(s, q), m, tf, fo);
                const double dil = 0.3 * fp->supersampling_factor * fp->supersampling_factor;
                cov = cov + vec3<double>{dil, 0.0, dil};
                const double det = cov.x * cov.z - cov.y * cov.y;
                if (det <= 0) continue;
                const vec2<double> pm = {(pr.x * 0.5 + 0.5) * f.up_w, (pr.y * 0.5 + 0.5) * f.up_h};
                ivec2_ bmin, bmax;
                if (!compute2DGaussianBounds3DGS(bmin, bmax, f.up_w, f.up_h, cov, pm)) continue;
                std::vector<double> sh(48);
                for (int k = 0; k < 48; ++k) sh[k] = p->sh48[48 * g + k];
                const vec3<double> c = p->sh_color(m, campos, sh.data());
                const double op = std::min((double)p->opac[g], 1.0);
                // the reference's dilog fit (random.cuh:88-104) evaluated in FP64
                const double cf[8] = {-6.09201442e-01, 1.79126616e+00, -2.14953223e+00, 1.26304372e+00,
                                      -4.59069895e-01, -1.87417414e-01, 1.99603130e+00, 5.99669467e-05};
                double dl = cf[0];
                for (int k = 1; k < 8; ++k) dl = dl * op + cf[k];
                if (1.0 - op > 0.0) dl += (1.0 - op) * std::log(1.0 - op);
                const double imp = 2.0 * M_PI * std::sqrt(det) * dl;
                const double vals[16] = {1, vm.z, pm.x, pm.y, cov.x, cov.y, cov.z, c.x, c.y, c.z, imp, pr.z,
                                         (double)bmin.x, (double)bmin.y, (double)bmax.x, (double)bmax.y};
                for (int k = 0; k < 16; ++k) o[k] = vals[k];
            }
        }
    });
    return GPS_OK;
}

// post_processing_kernels.cuh:143-189,191-218,222-285.  The neighbourhood is
// read from the un-blended input frame.
void orc_resolve_taa(const float* cur, const float* curDepth, const float* prev, const float* prevDepth,
                     const float M16[16], float blend, int32_t W, int32_t H, int32_t frameNumber, float* outc,
                     uint8_t* rgb8) {
    const mat4<float> M = mat4_from<float>(M16);
    auto C = [&](const float* b, int i) { return vec3<float>{b[3 * i], b[3 * i + 1], b[3 * i + 2]}; };
    for (int idx = 0; idx < W * H; ++idx) {
        const int x = idx % W, y = idx / W;
        vec3<float> color = C(cur, idx);
        const float depth = curDepth[idx];
        if (frameNumber > 0) {
            const vec4<float> clip = {(x + 0.5f) / W * 2.0f - 1.0f, (y + 0.5f) / H * 2.0f - 1.0f, depth, 1.0f};
            vec4<float> pp = M * clip;
            const float w = pp.w;
            pp = {pp.x / w, pp.y / w, pp.z / w, pp.w / w};
            const float px = (pp.x * 0.5f + 0.5f) * W - 0.5f;
            const float py = (pp.y * 0.5f + 0.5f) * H - 0.5f;
            if (px >= 0 && px < W && py >= 0 && py < H) {
                int x0 = gps::det_f2i_rz(std::floor(px)), y0 = gps::det_f2i_rz(std::floor(py));
                int x1 = x0 + 1, y1 = y0 + 1;
                x0 = (x0 < 0) ? 0 : ((x0 >= W) ? W - 1 : x0);
                y0 = (y0 < 0) ? 0 : ((y0 >= H) ? H - 1 : y0);
                x1 = (x1 < 0) ? 0 : ((x1 >= W) ? W - 1 : x1);
                y1 = (y1 < 0) ? 0 : ((y1 >= H) ? H - 1 : y1);
                const float tx = px - std::floor(px), ty = py - std::floor(py);
                const int i00 = y0 * W + x0, i10 = y0 * W + x1, i01 = y1 * W + x0, i11 = y1 * W + x1;
                auto mixv = [](vec3<float> a, vec3<float> b, float t) { return a * (1.0f - t) + b * t; };
                const vec3<float> top = mixv(C(prev, i00), C(prev, i10), tx);
                const vec3<float> bot = mixv(C(prev, i01), C(prev, i11), tx);
                const vec3<float> hist = mixv(top, bot, ty);
                (void)prevDepth;
                vec3<float> nMin = {1e9f, 1e9f, 1e9f}, nMax = {-1e9f, -1e9f, -1e9f};
                for (int dy = -1; dy <= 1; dy++) {
                    const int yy = y + dy;
                    if (yy < 0 || yy >= H) continue;
                    for (int dx = -1; dx <= 1; dx++) {
                        const int xx = x + dx;
                        if (xx < 0 || xx >= W) continue;
                        const vec3<float> c = C(cur, yy * W + xx);
                        nMin = vmin(nMin, c);
                        nMax = vmax(nMax, c);
                    }
                }
                const vec3<float> lo = nMin + (-0.05f), hi = nMax + 0.05f;
                const vec3<float> hc = {gclamp(hist.x, lo.x, hi.x), gclamp(hist.y, lo.y, hi.y), gclamp(hist.z, lo.z, hi.z)};
                color = mixv(color, hc, blend);
            }
        }
        outc[3 * idx] = color.x;
        outc[3 * idx + 1] = color.y;
        outc[3 * idx + 2] = color.z;
        const float o[3] = {color.x, color.y, color.z};
        for (int k = 0; k < 3; ++k) rgb8[3 * idx + k] = (uint8_t)std::round(gclamp(o[k], 0.0f, 1.0f) * 255.0f);  // glm::round
    }
}

// post_processing_kernels.cuh:108-140
void orc_resolve_progressive(float* color, float* accum, int32_t W, int32_t H, int32_t n, uint8_t* rgb8) {
    for (int idx = 0; idx < W * H; ++idx) {
        vec3<float> c = {color[3 * idx], color[3 * idx + 1], color[3 * idx + 2]};
        if (n > 1) c = c + vec3<float>{accum[3 * idx], accum[3 * idx + 1], accum[3 * idx + 2]};
        color[3 * idx] = c.x;
        color[3 * idx + 1] = c.y;
        color[3 * idx + 2] = c.z;
        accum[3 * idx] = accum[3 * idx + 1] = accum[3 * idx + 2] = 0.0f;
        c = c * (255.0f / static_cast<float>(n));
        const float o[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; ++k) {
            // glm::round = round half away from zero
            const float v = std::round(gclamp(o[k], 0.0f, 255.0f));
            rgb8[3 * idx + k] = (uint8_t)v;
        }
    }
}

void orc_rng_probe(uint32_t idx, uint32_t frame, int32_t draws, uint32_t* out) {
    uint2_ s = makeSeed(idx, frame);
    out[0] = s.x;
    out[1] = s.y;
    for (int i = 0; i < draws; ++i) {
        pcg2d(s);
        out[2 + 2 * i] = s.x;
        out[3 + 2 * i] = s.y;
    }
}
float orc_dilog(float x) { return dilog(x); }
float orc_inv_dilog(float x) { return inv_dilog(x); }
void orc_corrected_box_muller(float u1, float u2, float alpha, float* out2) {
    const float2_ r = correctedBoxMuller(u1, u2, alpha);
    out2[0] = r.x;
    out2[1] = r.y;
}
uint32_t orc_poisson(uint32_t s0, uint32_t s1, float lambda) {
    uint2_ s = {s0, s1};
    return poisson(s, lambda);
}
int32_t orc_stochastic_round(float x, float u) { return stochastic_round(x, u); }
int32_t orc_pos_to_memory(int32_t x, int32_t y, int32_t p, int32_t np, int32_t w, int32_t h) {
    return posToMemory(x, y, p, np, w, h);
}
uint64_t orc_depth_bits(float vz, float n, float f) { return convert_view_depth_bits(vz, n, f); }

void orc_detmath(int32_t kind, const float* in, float* out, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        const float x = in[i];
        float r = 0.0f;
        switch (kind) {
            case 0: r = gps::det_expf(x); break;
            case 1: r = gps::det_logf(x); break;
            case 2: r = gps::det_sinf(x); break;
            case 3: r = gps::det_cosf(x); break;
            case 4: r = std::sqrt(x); break;
            case 5: r = 1.0f / x; break;
            case 6: r = dilog(x); break;
            case 7: r = inv_dilog(x); break;
            case 8: r = (float)gps::det_f2i_rz(std::nearbyint(x)); break;
            case 9: r = boxMuller(x, in[(i + 1) % n]).x; break;
            case 10: r = correctedBoxMuller(x, in[(i + 1) % n], 0.7f).y; break;
            case 11: r = fp8e4m3_to_float((uint8_t)((uint32_t)(int)x & 0xFFu)); break;
        }
        out[i] = r;
    }
}

// computeColorFromSHFloat at view direction `dir` (campos = 0, pos = dir).
void orc_sh_eval(const float dir[3], const float* coeffs48, int32_t fp64, double out[3]) {
    if (fp64) {
        double c[48];
        for (int k = 0; k < 48; ++k) c[k] = coeffs48[k];
        const vec3<double> r = computeColorFromSHFloat(vec3<double>{dir[0], dir[1], dir[2]}, vec3<double>{0, 0, 0}, c);
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    } else {
        const vec3<float> r = computeColorFromSHFloat(vec3<float>{dir[0], dir[1], dir[2]}, vec3<float>{0, 0, 0}, coeffs48);
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    }
}

// computeColorFromCompactSH at view direction `dir`, decoded coefficients.
void orc_sh_eval_compact(const float dir[3], const float* coeffs48, int32_t fp64, double out[3]) {
    if (fp64) {
        double c[48];
        for (int k = 0; k < 48; ++k) c[k] = coeffs48[k];
        const vec3<double> r =
            computeColorFromCompactSH(vec3<double>{dir[0], dir[1], dir[2]}, vec3<double>{0, 0, 0}, c);
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    } else {
        const vec3<float> r =
            computeColorFromCompactSH(vec3<float>{dir[0], dir[1], dir[2]}, vec3<float>{0, 0, 0}, coeffs48);
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    }
}

}  // extern "C"
