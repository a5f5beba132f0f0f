// gps_bvh.hip — load-time Gaussian order of the reference's BVH reorder
// (GaussianBVH::build_bvh_and_reorder, src/core/rendering/occlusion/
// gaussian_bvh.cu:146-169), built on the GPU.
//
// The reference builds a tinybvh 1.6.7 binned-SAH BVH on the CPU
// (packages/tinybvh/tiny_bvh.h:2292-2420) over one flat triangle per Gaussian
// and then reorders the Gaussians by a depth-first walk of its leaves
// (gaussian_bvh.cu:14-34).  That walk pushes the leaves' primIdx *positions*
// (leftFirst + i), not the primitive ids stored there, and visits the right
// child first; reorder() (gaussian_bvh.cu:47-84) then moves Gaussian m to
// leafOrder[m].  So only the leaf ranges matter: a leaf holding positions
// [f, f + c) receives old indices [n - f - c, n - f) in order, i.e.
//     perm[q] = (n - f - c) + (q - f)      for q in [f, f + c).
// The leaf ranges depend only on the split decisions, and every quantity a
// decision reads (bin counts, bin bounds as min/max) is independent of the
// order of the Gaussians inside a node.  The build therefore runs
// level-synchronously: one pass bins every Gaussian of every open node,
// one thread per node evaluates the SAH exactly as tinybvh does, and a
// partition pass scatters each Gaussian to its child's range (ranks in
// LDS, one cursor atomic per node side and workgroup).  Nodes of at most 32 Gaussians are
// finished by one thread each with the sequential algorithm.
//
// Float min/max go through an order-preserving integer encoding (atomic
// min/max), which orders -0 below +0 where tinybvh keeps whichever zero came
// last; a zero-sign difference in a bound cannot change a split decision
// (the decision reads extents and areas, and a zero area makes the cost
// NaN, i.e. "split", for either sign).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/gps.h"
#include "gps_detmath.h"
#include "gps_project.h"

namespace {

constexpr int BINS = 8;                  // BVHBINS (tiny_bvh.h:104-106)
constexpr float FAR = 1e30f;             // BVH_FAR (tiny_bvh.h:144)
constexpr uint32_t MT_THRESHOLD = 50000; // MT_BUILD_THRESHOLD (tiny_bvh.h:148-150)
constexpr uint32_t TASKS = 256;          // task stack of BVH::Build (tiny_bvh.h:2308)
constexpr uint32_t SMALL = 32;           // nodes this small finish in one thread
constexpr uint32_t CHUNK = 512;          // positions per workgroup in the level passes
constexpr uint32_t THREADS = 256;
// Open nodes hold more than SMALL Gaussians, so a chunk overlaps at most
// ceil(CHUNK / (SMALL + 1)) + 1 of them.
constexpr uint32_t MAXN = (CHUNK + SMALL) / (SMALL + 1) + 1;
constexpr uint32_t BINW = 3 * BINS * 7;  // words of one node's bins
static_assert(MAXN == 17, "chunk/node bound");

// A Gaussian's fragment: tinybvh's min/max of the three triangle vertices.
// The triangle is flat in z, so bmax.z == bmin.z and is not stored.
struct alignas(8) Frag {
    float lx, ly, lz, hx, hy, pad;
};

struct ONode {  // an open node: a position range and its box
    float mn[3], mx[3];
    uint32_t first, count;
    uint32_t tc;     // task-stack depth tinybvh has when it splits this node
    uint32_t depth;  // tree depth, saturated at 6
};

struct Split {
    float lmn[3], lmx[3], rmn[3], rmx[3];
    float nmin, rpd;
    uint32_t axis, pos, left, kind;  // kind: 0 leaf, 1 split
    uint32_t curL, curR;             // partition cursors
};

__device__ __forceinline__ float fmin_(float a, float b) { return a < b ? a : b; }  // tiny_bvh.h:445
__device__ __forceinline__ float fmax_(float a, float b) { return a > b ? a : b; }  // tiny_bvh.h:446
__device__ __forceinline__ int32_t clampi(int32_t x, int32_t a, int32_t b) { return x > a ? (x < b ? x : b) : a; }

// x86-64 conversions (cvttss2si) as the reference host performs them.
__device__ __forceinline__ int32_t cvt_i32(float x) {
    return (x >= -2147483648.0f && x < 2147483648.0f) ? (int32_t)x : INT32_MIN;
}
__device__ __forceinline__ uint32_t cvt_u32(float x) {
    return (x >= -9223372036854775808.0f && x < 9223372036854775808.0f) ? (uint32_t)(uint64_t)(int64_t)x : 0u;
}

__device__ __forceinline__ uint32_t ford(float f) {
    const uint32_t u = gps::f2u_bits(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t e) {
    return gps::u2f_bits((e & 0x80000000u) ? (e & 0x7fffffffu) : ~e);
}

__device__ __forceinline__ float half_area(float x, float y, float z) {  // tiny_bvh.h:460
    return x < -FAR ? 0.0f : (x * y + y * z + z * x);
}

__device__ __forceinline__ float centroid(const Frag& f, uint32_t a) {
    return a == 0 ? (f.lx + f.hx) * 0.5f : a == 1 ? (f.ly + f.hy) * 0.5f : (f.lz + f.lz) * 0.5f;
}
__device__ __forceinline__ float fmin_axis(const Frag& f, int c) { return c == 0 ? f.lx : c == 1 ? f.ly : f.lz; }
__device__ __forceinline__ float fmax_axis(const Frag& f, int c) { return c == 0 ? f.hx : c == 1 ? f.hy : f.lz; }

// One axis of tinybvh's split search (tiny_bvh.h:2336-2369): prefix/suffix
// boxes and counts, then the first bin boundary with the lowest cost.
struct Best {
    float cost;
    uint32_t axis, pos;
    float lmn[3], lmx[3], rmn[3], rmx[3];
};
__device__ __forceinline__ void sah_axis(const uint32_t cnt[BINS], const float bmn[BINS][3], const float bmx[BINS][3],
                                         uint32_t a, Best& best) {
    float lBMin[BINS - 1][3], rBMin[BINS - 1][3], lBMax[BINS - 1][3], rBMax[BINS - 1][3];
    float ANL[BINS - 1], ANR[BINS - 1];
    float l1[3] = {FAR, FAR, FAR}, l2[3] = {-FAR, -FAR, -FAR}, r1[3] = {FAR, FAR, FAR}, r2[3] = {-FAR, -FAR, -FAR};
    uint32_t lN = 0, rN = 0;
#pragma unroll
    for (int i = 0; i < BINS - 1; ++i) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            lBMin[i][c] = l1[c] = fmin_(l1[c], bmn[i][c]);
            rBMin[BINS - 2 - i][c] = r1[c] = fmin_(r1[c], bmn[BINS - 1 - i][c]);
            lBMax[i][c] = l2[c] = fmax_(l2[c], bmx[i][c]);
            rBMax[BINS - 2 - i][c] = r2[c] = fmax_(r2[c], bmx[BINS - 1 - i][c]);
        }
        lN += cnt[i];
        rN += cnt[BINS - 1 - i];
        ANL[i] = lN == 0 ? FAR : half_area(l2[0] - l1[0], l2[1] - l1[1], l2[2] - l1[2]) * (float)lN;
        ANR[BINS - 2 - i] = rN == 0 ? FAR : half_area(r2[0] - r1[0], r2[1] - r1[1], r2[2] - r1[2]) * (float)rN;
    }
#pragma unroll
    for (int i = 0; i < BINS - 1; ++i) {
        const float C = ANL[i] + ANR[i];
        if (C < best.cost) {
            best.cost = C, best.axis = a, best.pos = (uint32_t)i;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                best.lmn[c] = lBMin[i][c], best.lmx[c] = lBMax[i][c];
                best.rmn[c] = rBMin[i][c], best.rmx[c] = rBMax[i][c];
            }
        }
    }
}

__device__ __forceinline__ float node_area(const float mn[3], const float mx[3]) {  // BVHBase::SA, tiny_bvh.h:8437-8441
    const float ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
    return ex * ey + ey * ez + ez * ex;
}

// Child task-stack depth (tiny_bvh.h:2393-2400): subtrees of nodes above
// depth 5 start on fresh threads (empty stack) in a threaded build; otherwise
// the left child runs with the right one pushed.
__device__ __forceinline__ uint32_t child_tc(uint32_t tc, uint32_t depth, bool threaded, bool left) {
    if (threaded && depth < 5) return 0u;
    return left ? tc + 1u : tc;
}

// ------------------------------------------------------------------ kernels
// gaussian_bvh.cu:150-155 + gaussian_bvh.cuh:20-26 + tiny_bvh.h:2256-2266.
__global__ __launch_bounds__(256) void bvh_frag_kernel(const float* __restrict__ means, const float* __restrict__ rots,
                                                       const float* __restrict__ scales, uint32_t n,
                                                       Frag* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float C[9];
    gps::scale_rot_to_cov(scales[3 * i], scales[3 * i + 1], scales[3 * i + 2], rots[4 * i], rots[4 * i + 1],
                          rots[4 * i + 2], rots[4 * i + 3], C);
    const float ex = 3.0f * gps::det_sqrt(C[0]), ey = 3.0f * gps::det_sqrt(C[4]), ez = 3.0f * gps::det_sqrt(C[8]);
    const float lx = means[3 * i] - ex, ly = means[3 * i + 1] - ey, lz = means[3 * i + 2] - ez;
    const float hx = means[3 * i] + ex, hy = means[3 * i + 1] + ey;
    // vertices (lx, ly, lz), (hx, ly, lz), (lx, hy, lz)
    Frag f;
    f.lx = fmin_(lx, fmin_(hx, lx));
    f.ly = fmin_(ly, fmin_(ly, hy));
    f.lz = lz;
    f.hx = fmax_(lx, fmax_(hx, lx));
    f.hy = fmax_(ly, fmax_(ly, hy));
    f.pad = 0.0f;
    out[i] = f;
}

// Root box: encoded min of bmin, max of bmax over all fragments.
__global__ __launch_bounds__(256) void bvh_root_kernel(const Frag* __restrict__ fr, uint32_t n, uint32_t* __restrict__ box) {
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const Frag f = fr[i];
        mn[0] = min(mn[0], ford(f.lx)), mn[1] = min(mn[1], ford(f.ly)), mn[2] = min(mn[2], ford(f.lz));
        mx[0] = max(mx[0], ford(f.hx)), mx[1] = max(mx[1], ford(f.hy)), mx[2] = max(mx[2], ford(f.lz));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        for (int m = 32; m >= 1; m >>= 1) {
            mn[c] = min(mn[c], (uint32_t)__shfl_xor((int)mn[c], m, 64));
            mx[c] = max(mx[c], (uint32_t)__shfl_xor((int)mx[c], m, 64));
        }
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 3; ++c) {
            atomicMin(&box[c], mn[c]);
            atomicMax(&box[3 + c], mx[c]);
        }
}

// Root node and the build constants (PrepareBuild, tiny_bvh.h:2253-2266;
// minDim, tiny_bvh.h:2310).
__global__ void bvh_root_node_kernel(const uint32_t* __restrict__ box, uint32_t n, ONode* __restrict__ root,
                                     float* __restrict__ min_dim) {
    if (threadIdx.x != 0) return;
    ONode r;
    for (int c = 0; c < 3; ++c) {
        r.mn[c] = fmin_(FAR, fdec(box[c]));
        r.mx[c] = fmax_(-FAR, fdec(box[3 + c]));
        min_dim[c] = (r.mx[c] - r.mn[c]) * 1e-20f;
    }
    r.first = 0, r.count = n, r.tc = 0, r.depth = 0;
    *root = r;
}

__global__ __launch_bounds__(256) void bvh_bins_init_kernel(uint32_t* __restrict__ bins, uint64_t words) {
    for (uint64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t w = (uint32_t)(i % BINW);
        bins[i] = w < 3 * BINS ? 0u : (w < 3 * BINS * 4 ? 0xffffffffu : 0u);
    }
}

// Open nodes overlapping positions [s, s + CHUNK): first index by binary
// search over the (position-ordered) node list, then up to MAXN nodes.
__device__ uint32_t chunk_nodes(const ONode* __restrict__ nodes, uint32_t nn, uint32_t s, uint32_t* k0) {
    uint32_t lo = 0, hi = nn;  // first node with first + count > s
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (nodes[mid].first + nodes[mid].count <= s)
            lo = mid + 1;
        else
            hi = mid;
    }
    uint32_t k = lo, m = 0;
    while (k + m < nn && m < MAXN && nodes[k + m].first < s + CHUNK) ++m;
    *k0 = k;
    return m;
}

__device__ __forceinline__ int find_slot(const uint32_t* s_first, const uint32_t* s_count, uint32_t m, uint32_t p) {
    for (uint32_t j = 0; j < m; ++j)
        if (p >= s_first[j] && p - s_first[j] < s_count[j]) return (int)j;
    return -1;
}

// Bins of every open node (tiny_bvh.h:2315-2331), LDS-aggregated per chunk.
__global__ __launch_bounds__(THREADS) void bvh_bin_kernel(const Frag* __restrict__ fr, uint32_t n,
                                                          const ONode* __restrict__ nodes, uint32_t nn,
                                                          uint32_t* __restrict__ bins) {
    __shared__ uint32_t s_k0, s_m;
    __shared__ uint32_t s_first[MAXN], s_count[MAXN];
    __shared__ float s_nmin[MAXN][3], s_rpd[MAXN][3];
    __shared__ uint32_t s_bins[MAXN * BINW];
    const uint32_t s = blockIdx.x * CHUNK, tid = threadIdx.x;
    if (tid == 0) {
        uint32_t k0;
        s_m = chunk_nodes(nodes, nn, s, &k0);
        s_k0 = k0;
    }
    __syncthreads();
    const uint32_t m = s_m;
    if (m == 0) return;  // uniform
    if (tid < m) {
        const ONode& nd = nodes[s_k0 + tid];
        s_first[tid] = nd.first;
        s_count[tid] = nd.count;
        for (int c = 0; c < 3; ++c) {
            s_nmin[tid][c] = nd.mn[c];
            s_rpd[tid][c] = (float)BINS / (nd.mx[c] - nd.mn[c]);
        }
    }
    for (uint32_t i = tid; i < m * BINW; i += THREADS) {
        const uint32_t w = i % BINW;
        s_bins[i] = w < 3 * BINS ? 0u : (w < 3 * BINS * 4 ? 0xffffffffu : 0u);
    }
    __syncthreads();
    for (uint32_t p = s + tid; p < s + CHUNK && p < n; p += THREADS) {
        const int j = find_slot(s_first, s_count, m, p);
        if (j < 0) continue;
        const Frag f = fr[p];
        uint32_t* b = s_bins + j * BINW;
        const uint32_t emn[3] = {ford(f.lx), ford(f.ly), ford(f.lz)};
        const uint32_t emx[3] = {ford(f.hx), ford(f.hy), ford(f.lz)};
#pragma unroll
        for (uint32_t a = 0; a < 3; ++a) {
            const int32_t bi = clampi(cvt_i32((centroid(f, a) - s_nmin[j][a]) * s_rpd[j][a]), 0, BINS - 1);
            const uint32_t o = a * BINS + (uint32_t)bi;
            atomicAdd(&b[o], 1u);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                atomicMin(&b[3 * BINS + 3 * o + c], emn[c]);
                atomicMax(&b[3 * BINS * 4 + 3 * o + c], emx[c]);
            }
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < m * BINW; i += THREADS) {
        const uint32_t j = i / BINW, w = i % BINW, v = s_bins[i];
        uint32_t* g = bins + (uint64_t)(s_k0 + j) * BINW + w;
        if (w < 3 * BINS) {
            if (v) atomicAdd(g, v);
        } else if (w < 3 * BINS * 4) {
            if (v != 0xffffffffu) atomicMin(g, v);
        } else if (v) {
            atomicMax(g, v);
        }
    }
}

// One thread per open node: tinybvh's decision (tiny_bvh.h:2333-2391) from
// the node's bins.  Writes the split record and the number of children that
// stay open (nb) and that finish in one thread (ns).
__global__ __launch_bounds__(256) void bvh_split_kernel(const ONode* __restrict__ nodes, uint32_t nn,
                                                        const uint32_t* __restrict__ bins,
                                                        const float* __restrict__ min_dim, int threaded,
                                                        Split* __restrict__ splits, uint2* __restrict__ counts,
                                                        uint32_t* __restrict__ err) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const ONode nd = nodes[k];
    const uint32_t* b = bins + (uint64_t)k * BINW;
    Best best;
    best.cost = FAR, best.axis = 0, best.pos = 0;
    for (int c = 0; c < 3; ++c) best.lmn[c] = best.lmx[c] = best.rmn[c] = best.rmx[c] = 0.0f;
    for (uint32_t a = 0; a < 3; ++a) {
        if (!((nd.mx[a] - nd.mn[a]) > min_dim[a])) continue;
        uint32_t cnt[BINS];
        float bmn[BINS][3], bmx[BINS][3];
#pragma unroll
        for (int i = 0; i < BINS; ++i) {
            const uint32_t o = a * BINS + (uint32_t)i;
            cnt[i] = b[o];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                bmn[i][c] = cnt[i] ? fmin_(FAR, fdec(b[3 * BINS + 3 * o + c])) : FAR;
                bmx[i][c] = cnt[i] ? fmax_(-FAR, fdec(b[3 * BINS * 4 + 3 * o + c])) : -FAR;
            }
        }
        sah_axis(cnt, bmn, bmx, a, best);
    }
    const float rSAV = 1.0f / node_area(nd.mn, nd.mx);
    const float splitCost = 1.0f + 1.0f * rSAV * best.cost;
    const float noSplitCost = (float)nd.count * 1.0f;
    Split sp;
    for (int c = 0; c < 3; ++c) {
        sp.lmn[c] = best.lmn[c], sp.lmx[c] = best.lmx[c];
        sp.rmn[c] = best.rmn[c], sp.rmx[c] = best.rmx[c];
    }
    sp.axis = best.axis, sp.pos = best.pos;
    sp.nmin = nd.mn[best.axis];
    sp.rpd = (float)BINS / (nd.mx[best.axis] - nd.mn[best.axis]);
    uint32_t left = 0;
    for (uint32_t i = 0; i <= best.pos; ++i) left += b[best.axis * BINS + i];
    sp.left = left;
    sp.curL = 0, sp.curR = 0;
    const uint32_t right = nd.count - left;
    bool split = !(splitCost >= noSplitCost) && left != 0 && right != 0 && nd.tc != TASKS;
    // With no candidate on any axis tinybvh may still split on axis 0 when
    // the node area exceeds ~1e30 / count; its partition then converts
    // out-of-range bin values differently from the binning, so the counts
    // above do not apply.  Such scenes are reported, not built.
    if (split && !((nd.mx[best.axis] - nd.mn[best.axis]) > min_dim[best.axis])) {
        atomicOr(err, 1u);
        split = false;
    }
    sp.kind = split ? 1u : 0u;
    splits[k] = sp;
    uint2 c2 = make_uint2(0u, 0u);
    if (split) {
        c2.x = (left > SMALL ? 1u : 0u) + (right > SMALL ? 1u : 0u);
        c2.y = 2u - c2.x;
    }
    counts[k] = c2;
}

// Exclusive scan of uint2 pairs: per-block sums, a one-workgroup scan of the
// sums, per-block scan plus offset.  1024 elements per block.
constexpr uint32_t SCAN_BLOCK = 1024;
__device__ __forceinline__ uint2 add2(uint2 a, uint2 b) { return make_uint2(a.x + b.x, a.y + b.y); }

__device__ uint2 block_exclusive(uint2 v, uint2* s_tmp, uint2* total) {
    // 256 threads; wave scans then a scan of the four wave totals.
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint2 x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t tx = (uint32_t)__shfl_up((int)x.x, d, 64), ty = (uint32_t)__shfl_up((int)x.y, d, 64);
        if ((int)lane >= d) x.x += tx, x.y += ty;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint2 off = make_uint2(0u, 0u);
    for (uint32_t w = 0; w < wave; ++w) off = add2(off, s_tmp[w]);
    *total = add2(add2(s_tmp[0], s_tmp[1]), add2(s_tmp[2], s_tmp[3]));
    __syncthreads();
    return make_uint2(off.x + x.x - v.x, off.y + x.y - v.y);
}

__global__ __launch_bounds__(256) void scan_sums_kernel(const uint2* __restrict__ in, uint32_t n, uint2* __restrict__ sums) {
    __shared__ uint2 s_tmp[4];
    const uint32_t base = blockIdx.x * SCAN_BLOCK;
    uint2 v = make_uint2(0u, 0u);
    for (uint32_t i = 0; i < SCAN_BLOCK / 256; ++i) {
        const uint32_t e = base + i * 256 + threadIdx.x;
        if (e < n) v = add2(v, in[e]);
    }
    uint2 total;
    (void)block_exclusive(v, s_tmp, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void scan_top_kernel(uint2* __restrict__ sums, uint32_t nb, uint2* __restrict__ grand) {
    __shared__ uint2 s_tmp[4];
    uint2 carry = make_uint2(0u, 0u);
    for (uint32_t b = 0; b < nb; b += 256) {
        const uint32_t e = b + threadIdx.x;
        const uint2 v = e < nb ? sums[e] : make_uint2(0u, 0u);
        uint2 total;
        const uint2 ex = block_exclusive(v, s_tmp, &total);
        if (e < nb) sums[e] = add2(carry, ex);
        carry = add2(carry, total);
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ __launch_bounds__(256) void scan_apply_kernel(uint2* __restrict__ data, uint32_t n, const uint2* __restrict__ sums) {
    __shared__ uint2 s_tmp[4];
    const uint32_t base = blockIdx.x * SCAN_BLOCK;
    uint2 carry = sums[blockIdx.x];
    for (uint32_t i = 0; i < SCAN_BLOCK / 256; ++i) {
        const uint32_t e = base + i * 256 + threadIdx.x;
        const uint2 v = e < n ? data[e] : make_uint2(0u, 0u);
        uint2 total;
        const uint2 ex = block_exclusive(v, s_tmp, &total);
        if (e < n) data[e] = add2(carry, ex);
        carry = add2(carry, total);
    }
}

// Children of split nodes into the next open list and the small list, in
// position order (left before right, nodes in order).
__global__ __launch_bounds__(256) void bvh_emit_kernel(const ONode* __restrict__ nodes, uint32_t nn,
                                                       const Split* __restrict__ splits, const uint2* __restrict__ offs,
                                                       int threaded, ONode* __restrict__ next, ONode* __restrict__ small) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const Split sp = splits[k];
    if (!sp.kind) return;
    const ONode nd = nodes[k];
    uint2 o = offs[k];
    ONode L, R;
    for (int c = 0; c < 3; ++c) {
        L.mn[c] = sp.lmn[c], L.mx[c] = sp.lmx[c];
        R.mn[c] = sp.rmn[c], R.mx[c] = sp.rmx[c];
    }
    L.first = nd.first, L.count = sp.left;
    R.first = nd.first + sp.left, R.count = nd.count - sp.left;
    L.tc = child_tc(nd.tc, nd.depth, threaded != 0, true);
    R.tc = child_tc(nd.tc, nd.depth, threaded != 0, false);
    L.depth = R.depth = nd.depth < 6 ? nd.depth + 1 : 6;
    if (L.count > SMALL) next[o.x++] = L; else small[o.y++] = L;
    if (R.count > SMALL) next[o.x++] = R; else small[o.y++] = R;
}

// Moves each Gaussian of a split node into its child's range; leaves write
// their final indices (tiny_bvh.h:2372-2380 with any order inside a range).
__global__ __launch_bounds__(THREADS) void bvh_partition_kernel(const Frag* __restrict__ src, Frag* __restrict__ dst,
                                                                uint32_t n, const ONode* __restrict__ nodes,
                                                                uint32_t nn, Split* __restrict__ splits,
                                                                uint32_t* __restrict__ perm, uint32_t* __restrict__ err) {
    __shared__ uint32_t s_k0, s_m;
    __shared__ uint32_t s_first[MAXN], s_count[MAXN], s_axis[MAXN], s_pos[MAXN], s_left[MAXN], s_kind[MAXN];
    __shared__ float s_nmin[MAXN], s_rpd[MAXN];
    __shared__ uint32_t s_cnt[2 * MAXN], s_base[2 * MAXN];
    const uint32_t s = blockIdx.x * CHUNK, tid = threadIdx.x;
    if (tid == 0) {
        uint32_t k0;
        s_m = chunk_nodes(nodes, nn, s, &k0);
        s_k0 = k0;
    }
    __syncthreads();
    const uint32_t m = s_m;
    if (m == 0) return;  // uniform
    if (tid < m) {
        const ONode& nd = nodes[s_k0 + tid];
        const Split& sp = splits[s_k0 + tid];
        s_first[tid] = nd.first, s_count[tid] = nd.count;
        s_axis[tid] = sp.axis, s_pos[tid] = sp.pos, s_left[tid] = sp.left, s_kind[tid] = sp.kind;
        s_nmin[tid] = sp.nmin, s_rpd[tid] = sp.rpd;
    }
    for (uint32_t i = tid; i < 2 * MAXN; i += THREADS) s_cnt[i] = 0u;
    __syncthreads();
    // ranks inside the workgroup per (node, side), then one global cursor
    // bump per (node, side) for the whole workgroup
    constexpr uint32_t R = CHUNK / THREADS;
    int key[R];
    uint32_t rank[R];
    Frag f[R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        const uint32_t p = s + r * THREADS + tid;
        const int j = p < n ? find_slot(s_first, s_count, m, p) : -1;
        key[r] = -1;
        if (j < 0) continue;
        if (!s_kind[j]) {
            // leaf: positions [f, f + c) take old indices n - f - c + (q - f)
            perm[p] = (n - s_first[j] - s_count[j]) + (p - s_first[j]);
            continue;
        }
        f[r] = src[p];
        const uint32_t a = s_axis[j];
        const float c = centroid(f[r], a);
        const int32_t bi = clampi((int32_t)cvt_u32((c - s_nmin[j]) * s_rpd[j]), 0, BINS - 1);
        key[r] = 2 * j + ((uint32_t)bi <= s_pos[j] ? 0 : 1);
        rank[r] = atomicAdd(&s_cnt[key[r]], 1u);
    }
    __syncthreads();
    if (tid < 2 * m && s_cnt[tid]) {
        Split& sp = splits[s_k0 + (tid >> 1)];
        s_base[tid] = atomicAdd((tid & 1) ? &sp.curR : &sp.curL, s_cnt[tid]);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
        if (key[r] < 0) continue;
        const uint32_t jj = (uint32_t)(key[r] >> 1);
        const uint32_t slot = s_base[key[r]] + rank[r];
        const uint32_t cap = (key[r] & 1) ? s_count[jj] - s_left[jj] : s_left[jj];
        if (slot < cap)
            dst[s_first[jj] + ((key[r] & 1) ? s_left[jj] + slot : slot)] = f[r];
        else
            atomicOr(err, 2u);
    }
}

// Nodes of at most SMALL Gaussians: the sequential builder of
// tiny_bvh.h:2292-2420 on the node's range, one thread per node, with the
// same decisions as the level passes.  Leaves write their final indices.
struct Task {
    float mn[3], mx[3];
    uint32_t first, count, tc, depth;
};
__global__ __launch_bounds__(64) void bvh_small_kernel(Frag* __restrict__ fr, uint32_t n,
                                                       const ONode* __restrict__ small, uint32_t ns,
                                                       const float* __restrict__ min_dim, int threaded,
                                                       uint32_t* __restrict__ perm) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const ONode sn = small[k];
    const uint32_t base = sn.first;  // this thread owns [base, base + count)
    Task stack[SMALL];
    uint32_t sp = 0;
    Task t;
    for (int c = 0; c < 3; ++c) t.mn[c] = sn.mn[c], t.mx[c] = sn.mx[c];
    t.first = 0, t.count = sn.count, t.tc = sn.tc, t.depth = sn.depth;
    const float md[3] = {min_dim[0], min_dim[1], min_dim[2]};
    while (true) {
        bool leaf = true;
        Best best;
        best.cost = FAR, best.axis = 0, best.pos = 0;
        for (int c = 0; c < 3; ++c) best.lmn[c] = best.lmx[c] = best.rmn[c] = best.rmx[c] = 0.0f;
        float rpd3[3];
        for (int c = 0; c < 3; ++c) rpd3[c] = (float)BINS / (t.mx[c] - t.mn[c]);
        if (t.count > 1) {
            for (uint32_t a = 0; a < 3; ++a) {
                if (!((t.mx[a] - t.mn[a]) > md[a])) continue;
                uint32_t cnt[BINS];
                float bmn[BINS][3], bmx[BINS][3];
#pragma unroll
                for (int i = 0; i < BINS; ++i) {
                    cnt[i] = 0;
#pragma unroll
                    for (int c = 0; c < 3; ++c) bmn[i][c] = FAR, bmx[i][c] = -FAR;
                }
                for (uint32_t i = 0; i < t.count; ++i) {
                    const Frag f = fr[base + t.first + i];
                    const int32_t bi = clampi(cvt_i32((centroid(f, a) - t.mn[a]) * rpd3[a]), 0, BINS - 1);
#pragma unroll
                    for (int b = 0; b < BINS; ++b)
                        if (bi == b) {
                            ++cnt[b];
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                bmn[b][c] = fmin_(bmn[b][c], fmin_axis(f, c));
                                bmx[b][c] = fmax_(bmx[b][c], fmax_axis(f, c));
                            }
                        }
                }
                sah_axis(cnt, bmn, bmx, a, best);
            }
            const float rSAV = 1.0f / node_area(t.mn, t.mx);
            const float splitCost = 1.0f + 1.0f * rSAV * best.cost;
            if (!(splitCost >= (float)t.count * 1.0f)) {
                // in-place partition of the node's fragments (tiny_bvh.h:2372-2380)
                const uint32_t a = best.axis;
                const float rpd = rpd3[a], nmin = t.mn[a];
                uint32_t j = t.first + t.count, src = t.first;
                for (uint32_t i = 0; i < t.count; ++i) {
                    const Frag f = fr[base + src];
                    const int32_t bi = clampi((int32_t)cvt_u32((centroid(f, a) - nmin) * rpd), 0, BINS - 1);
                    if ((uint32_t)bi <= best.pos) {
                        ++src;
                    } else {
                        --j;
                        fr[base + src] = fr[base + j];
                        fr[base + j] = f;
                    }
                }
                const uint32_t left = src - t.first, right = t.count - left;
                if (left != 0 && right != 0 && t.tc != TASKS) {
                    leaf = false;
                    Task L, R;
                    for (int c = 0; c < 3; ++c) {
                        L.mn[c] = best.lmn[c], L.mx[c] = best.lmx[c];
                        R.mn[c] = best.rmn[c], R.mx[c] = best.rmx[c];
                    }
                    L.first = t.first, L.count = left;
                    R.first = t.first + left, R.count = right;
                    L.tc = child_tc(t.tc, t.depth, threaded != 0, true);
                    R.tc = child_tc(t.tc, t.depth, threaded != 0, false);
                    L.depth = R.depth = t.depth < 6 ? t.depth + 1 : 6;
                    stack[sp++] = R;  // at most one pending right per level, depth <= SMALL
                    t = L;
                }
            }
        }
        if (leaf) {
            const uint32_t f0 = base + t.first;
            for (uint32_t q = f0; q < f0 + t.count; ++q) perm[q] = (n - f0 - t.count) + (q - f0);
            if (sp == 0) break;
            t = stack[--sp];
        }
    }
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" gps_status gps_bvh_order(const float* means, const float* rotations, const float* scales, uint64_t n,
                                    uint32_t* perm) {
    if ((!means || !rotations || !scales || !perm) && n) return GPS_ERR_INVALID_ARG;
    if (n == 0) return GPS_OK;
    if (n > 0x7fffffffull) return GPS_ERR_UNSUPPORTED;
    const uint32_t N = (uint32_t)n;
    const uint32_t maxOpen = N / (SMALL + 1) + 1;  // open nodes are disjoint and hold > SMALL each
    const uint32_t maxSmall = 2 * maxOpen + 1;
    const uint32_t nchunks = (N + CHUNK - 1) / CHUNK;
    const uint32_t maxScanBlocks = (maxOpen + SCAN_BLOCK - 1) / SCAN_BLOCK;
    DevBuf bFrag[2], bPerm, bNodes[2], bSmall, bBins, bSplit, bCnt, bSums, bMisc, bIn;
#define GPS_BVH_HIP(call)                              \
    do {                                               \
        if ((call) != hipSuccess) return GPS_ERR_HIP;  \
    } while (0)
    GPS_BVH_HIP(hipMalloc(&bFrag[0].p, (uint64_t)N * sizeof(Frag)));
    GPS_BVH_HIP(hipMalloc(&bFrag[1].p, (uint64_t)N * sizeof(Frag)));
    GPS_BVH_HIP(hipMalloc(&bPerm.p, (uint64_t)N * 4));
    GPS_BVH_HIP(hipMalloc(&bNodes[0].p, (uint64_t)maxOpen * sizeof(ONode)));
    GPS_BVH_HIP(hipMalloc(&bNodes[1].p, (uint64_t)maxOpen * sizeof(ONode)));
    GPS_BVH_HIP(hipMalloc(&bSmall.p, (uint64_t)maxSmall * sizeof(ONode)));
    GPS_BVH_HIP(hipMalloc(&bBins.p, (uint64_t)maxOpen * BINW * 4));
    GPS_BVH_HIP(hipMalloc(&bSplit.p, (uint64_t)maxOpen * sizeof(Split)));
    GPS_BVH_HIP(hipMalloc(&bCnt.p, (uint64_t)maxOpen * sizeof(uint2)));
    GPS_BVH_HIP(hipMalloc(&bSums.p, (uint64_t)(maxScanBlocks + 1) * sizeof(uint2)));
    GPS_BVH_HIP(hipMalloc(&bMisc.p, 256));
    Frag* frag[2] = {(Frag*)bFrag[0].p, (Frag*)bFrag[1].p};
    ONode* open[2] = {(ONode*)bNodes[0].p, (ONode*)bNodes[1].p};
    ONode* small = (ONode*)bSmall.p;
    uint32_t* bins = (uint32_t*)bBins.p;
    Split* splits = (Split*)bSplit.p;
    uint2* cnt = (uint2*)bCnt.p;
    uint2* sums = (uint2*)bSums.p;
    uint32_t* box = (uint32_t*)bMisc.p;                  // 6 words
    float* min_dim = (float*)((char*)bMisc.p + 64);      // 3 floats
    uint2* grand = (uint2*)((char*)bMisc.p + 128);       // 1 pair
    uint32_t* err = (uint32_t*)((char*)bMisc.p + 192);   // 1 word
    uint32_t* dperm = (uint32_t*)bPerm.p;
    const int threaded = N >= MT_THRESHOLD ? 1 : 0;

    // fragments, uploaded in slices
    const uint32_t SLICE = 1u << 23;
    GPS_BVH_HIP(hipMalloc(&bIn.p, (uint64_t)(N < SLICE ? N : SLICE) * 40));
    for (uint32_t s0 = 0; s0 < N; s0 += SLICE) {
        const uint32_t c = N - s0 < SLICE ? N - s0 : SLICE;
        float* dm = (float*)bIn.p;
        float* dr = dm + 3ull * c;
        float* ds = dr + 4ull * c;
        GPS_BVH_HIP(hipMemcpy(dm, means + 3ull * s0, 12ull * c, hipMemcpyHostToDevice));
        GPS_BVH_HIP(hipMemcpy(dr, rotations + 4ull * s0, 16ull * c, hipMemcpyHostToDevice));
        GPS_BVH_HIP(hipMemcpy(ds, scales + 3ull * s0, 12ull * c, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(bvh_frag_kernel, dim3((c + 255) / 256), dim3(256), 0, 0, dm, dr, ds, c, frag[0] + s0);
        GPS_BVH_HIP(hipGetLastError());
    }
    GPS_BVH_HIP(hipDeviceSynchronize());
    (void)hipFree(bIn.p);
    bIn.p = nullptr;

    const uint32_t boxInit[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    GPS_BVH_HIP(hipMemcpy(box, boxInit, sizeof(boxInit), hipMemcpyHostToDevice));
    GPS_BVH_HIP(hipMemset(err, 0, 4));
    hipLaunchKernelGGL(bvh_root_kernel, dim3(1024), dim3(256), 0, 0, frag[0], N, box);
    hipLaunchKernelGGL(bvh_root_node_kernel, dim3(1), dim3(64), 0, 0, box, N, open[0], min_dim);
    GPS_BVH_HIP(hipGetLastError());

    int cur = 0, lv = 0;
    uint32_t nopen = 1;
    if (N <= SMALL) {
        GPS_BVH_HIP(hipMemcpy(small, open[0], sizeof(ONode), hipMemcpyDeviceToDevice));
        hipLaunchKernelGGL(bvh_small_kernel, dim3(1), dim3(64), 0, 0, frag[0], N, small, 1u, min_dim, threaded, dperm);
        GPS_BVH_HIP(hipGetLastError());
        nopen = 0;
    }
    while (nopen > 0) {
        const uint64_t words = (uint64_t)nopen * BINW;
        hipLaunchKernelGGL(bvh_bins_init_kernel, dim3((uint32_t)((words + 255) / 256 < 65535 ? (words + 255) / 256 : 65535)),
                           dim3(256), 0, 0, bins, words);
        hipLaunchKernelGGL(bvh_bin_kernel, dim3(nchunks), dim3(THREADS), 0, 0, frag[cur], N, open[lv], nopen, bins);
        hipLaunchKernelGGL(bvh_split_kernel, dim3((nopen + 255) / 256), dim3(256), 0, 0, open[lv], nopen, bins,
                           min_dim, threaded, splits, cnt, err);
        const uint32_t nsb = (nopen + SCAN_BLOCK - 1) / SCAN_BLOCK;
        hipLaunchKernelGGL(scan_sums_kernel, dim3(nsb), dim3(256), 0, 0, cnt, nopen, sums);
        hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(256), 0, 0, sums, nsb, grand);
        hipLaunchKernelGGL(scan_apply_kernel, dim3(nsb), dim3(256), 0, 0, cnt, nopen, sums);
        hipLaunchKernelGGL(bvh_emit_kernel, dim3((nopen + 255) / 256), dim3(256), 0, 0, open[lv], nopen, splits, cnt,
                           threaded, open[lv ^ 1], small);
        hipLaunchKernelGGL(bvh_partition_kernel, dim3(nchunks), dim3(THREADS), 0, 0, frag[cur], frag[cur ^ 1], N,
                           open[lv], nopen, splits, dperm, err);
        GPS_BVH_HIP(hipGetLastError());
        uint2 g;
        GPS_BVH_HIP(hipMemcpy(&g, grand, sizeof(g), hipMemcpyDeviceToHost));
        if (g.x > maxOpen || g.y > maxSmall) return GPS_ERR_HIP;  // cannot happen: children are disjoint ranges
        if (g.y) {
            hipLaunchKernelGGL(bvh_small_kernel, dim3((g.y + 63) / 64), dim3(64), 0, 0, frag[cur ^ 1], N, small, g.y,
                               min_dim, threaded, dperm);
            GPS_BVH_HIP(hipGetLastError());
        }
        cur ^= 1;
        lv ^= 1;
        nopen = g.x;
    }
    uint32_t herr = 0;
    GPS_BVH_HIP(hipMemcpy(&herr, err, 4, hipMemcpyDeviceToHost));
    if (herr) return GPS_ERR_UNSUPPORTED;
    GPS_BVH_HIP(hipMemcpy(perm, dperm, 4ull * N, hipMemcpyDeviceToHost));
#undef GPS_BVH_HIP
    return GPS_OK;
}
