// orc_bvh.cpp — CPU oracle of the load-time Gaussian order the reference
// gives a scene that is neither Morton-sorted nor marked "bvh_sorted_order"
// (GaussianBVH::build_bvh_and_reorder,
// src/core/rendering/occlusion/gaussian_bvh.cu:146-169, called from
// gaussian_point_splatting.cu:791-797).
//
// TEST INFRASTRUCTURE.  The checker of gps_bvh_order; the product library
// never links or loads it.
//
// Steps, each as the reference performs it:
//   fragments   gaussian_bvh.cu:150-155: 3-sigma box of the covariance
//               diagonal (gaussian_types.cuh:38-43), handed to tinybvh as the
//               triangle (min), (max.x, min.y, min.z), (min.x, max.y, min.z)
//               (gaussian_bvh.cuh:20-26); tinybvh keeps the vertex min/max
//               (tiny_bvh.h:2256-2266), so every fragment is flat in z.
//   build       tinybvh 1.6.7 BVH::Build(uint32_t, uint32_t),
//               packages/tinybvh/tiny_bvh.h:2292-2420: 8 bins per axis,
//               c_int = c_trav = 1, in-place partition, 256-entry task stack,
//               subtrees above depth 5 on their own threads when the scene
//               has at least 50000 Gaussians (MT_BUILD_THRESHOLD, :148-150).
//               Threads change node numbering only, so they run in turn here.
//   leaf order  GaussianBVH::collect_leaf_order (gaussian_bvh.cu:14-34):
//               depth-first with the right child popped first, pushing the
//               leaf's primIdx *positions* (leftFirst + i), not the primitive
//               ids stored there.
//   reorder     GaussianBVH::reorder (gaussian_bvh.cu:47-84, REORDER_SIMPLE
//               undefined): the cycle walk leaves new[leafOrder[m]] = old[m].
// Float-to-integer conversions follow x86-64 (cvttss2si): out of range or
// NaN gives INT32_MIN for int, and the low 32 bits of INT64_MIN (0) for the
// unsigned conversion of the partition.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "orc_ref.h"

using namespace orc;

namespace {

constexpr int BINS = 8;              // BVHBINS, tiny_bvh.h:104-106
constexpr float FAR = 1e30f;         // BVH_FAR, tiny_bvh.h:144
constexpr uint32_t MT_THRESHOLD = 50000u;

inline float fmin_(float a, float b) { return a < b ? a : b; }  // tinybvh_min, tiny_bvh.h:445
inline float fmax_(float a, float b) { return a > b ? a : b; }  // tinybvh_max, tiny_bvh.h:446
inline int32_t clampi(int32_t x, int32_t a, int32_t b) { return x > a ? (x < b ? x : b) : a; }

inline int32_t cvt_i32(float x) {
    if (!(x >= -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
    return (int32_t)x;
}
inline uint32_t cvt_u32(float x) {
    if (!(x >= -9223372036854775808.0f && x < 9223372036854775808.0f)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)x;
}

struct V3 {
    float v[3];
};
inline V3 vmin3(const V3& a, const V3& b) { return {{fmin_(a.v[0], b.v[0]), fmin_(a.v[1], b.v[1]), fmin_(a.v[2], b.v[2])}}; }
inline V3 vmax3(const V3& a, const V3& b) { return {{fmax_(a.v[0], b.v[0]), fmax_(a.v[1], b.v[1]), fmax_(a.v[2], b.v[2])}}; }
// tinybvh_half_area, tiny_bvh.h:460
inline float half_area(const V3& e) {
    return e.v[0] < -FAR ? 0.0f : (e.v[0] * e.v[1] + e.v[1] * e.v[2] + e.v[2] * e.v[0]);
}
inline V3 sub3(const V3& a, const V3& b) { return {{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}}; }

struct Frag {
    V3 bmin, bmax;
};
struct Node {
    V3 mn, mx;
    uint32_t leftFirst, triCount;
};

struct Builder {
    std::vector<Frag> frag;
    std::vector<uint32_t> prim;
    std::vector<Node> node;
    uint32_t next = 2;  // node 1 stays unused (tiny_bvh.h:2287)
    bool threaded = false;
    V3 minDim;

    // tiny_bvh.h:2292-2420
    void build(uint32_t ni, uint32_t depth) {
        uint32_t task[256], taskCount = 0;
        while (true) {
            while (true) {
                Node& nd = node[ni];
                V3 binMin[3][BINS], binMax[3][BINS];
                uint32_t count[3][BINS] = {};
                for (int a = 0; a < 3; ++a)
                    for (int i = 0; i < BINS; ++i) binMin[a][i] = {{FAR, FAR, FAR}}, binMax[a][i] = {{-FAR, -FAR, -FAR}};
                V3 rpd3, nmin3 = nd.mn;
                for (int a = 0; a < 3; ++a) rpd3.v[a] = (float)BINS / (nd.mx.v[a] - nd.mn.v[a]);
                for (uint32_t i = 0; i < nd.triCount; ++i) {
                    const Frag& f = frag[prim[nd.leftFirst + i]];
                    for (int a = 0; a < 3; ++a) {
                        const float c = (f.bmin.v[a] + f.bmax.v[a]) * 0.5f;
                        const int32_t b = clampi(cvt_i32((c - nmin3.v[a]) * rpd3.v[a]), 0, BINS - 1);
                        binMin[a][b] = vmin3(binMin[a][b], f.bmin);
                        binMax[a][b] = vmax3(binMax[a][b], f.bmax);
                        ++count[a][b];
                    }
                }
                float splitCost = FAR;
                const float sa = half_area(sub3(nd.mx, nd.mn));  // BVHBase::SA, tiny_bvh.h:8437-8441
                const float rSAV = 1.0f / sa;
                uint32_t bestAxis = 0, bestPos = 0;
                V3 bLMin{{0, 0, 0}}, bLMax{{0, 0, 0}}, bRMin{{0, 0, 0}}, bRMax{{0, 0, 0}};
                for (int a = 0; a < 3; ++a) {
                    if (!((nd.mx.v[a] - nd.mn.v[a]) > minDim.v[a])) continue;
                    V3 lBMin[BINS - 1], rBMin[BINS - 1], lBMax[BINS - 1], rBMax[BINS - 1];
                    V3 l1{{FAR, FAR, FAR}}, l2{{-FAR, -FAR, -FAR}}, r1{{FAR, FAR, FAR}}, r2{{-FAR, -FAR, -FAR}};
                    float ANL[BINS - 1], ANR[BINS - 1];
                    uint32_t lN = 0, rN = 0;
                    for (int i = 0; i < BINS - 1; ++i) {
                        lBMin[i] = l1 = vmin3(l1, binMin[a][i]);
                        rBMin[BINS - 2 - i] = r1 = vmin3(r1, binMin[a][BINS - 1 - i]);
                        lBMax[i] = l2 = vmax3(l2, binMax[a][i]);
                        rBMax[BINS - 2 - i] = r2 = vmax3(r2, binMax[a][BINS - 1 - i]);
                        lN += count[a][i];
                        rN += count[a][BINS - 1 - i];
                        ANL[i] = lN == 0 ? FAR : half_area(sub3(l2, l1)) * (float)lN;
                        ANR[BINS - 2 - i] = rN == 0 ? FAR : half_area(sub3(r2, r1)) * (float)rN;
                    }
                    for (int i = 0; i < BINS - 1; ++i) {
                        const float C = ANL[i] + ANR[i];
                        if (C < splitCost) {
                            splitCost = C, bestAxis = (uint32_t)a, bestPos = (uint32_t)i;
                            bLMin = lBMin[i], bRMin = rBMin[i], bLMax = lBMax[i], bRMax = rBMax[i];
                        }
                    }
                }
                splitCost = 1.0f + 1.0f * rSAV * splitCost;
                const float noSplitCost = (float)nd.triCount * 1.0f;
                if (splitCost >= noSplitCost) break;
                uint32_t j = nd.leftFirst + nd.triCount, src = nd.leftFirst;
                const float rpd = rpd3.v[bestAxis], nmin = nmin3.v[bestAxis];
                for (uint32_t i = 0; i < nd.triCount; ++i) {
                    const Frag& f = frag[prim[src]];
                    int32_t b = (int32_t)cvt_u32(((f.bmin.v[bestAxis] + f.bmax.v[bestAxis]) * 0.5f - nmin) * rpd);
                    b = clampi(b, 0, BINS - 1);
                    if ((uint32_t)b <= bestPos)
                        ++src;
                    else
                        std::swap(prim[src], prim[--j]);
                }
                const uint32_t leftCount = src - nd.leftFirst, rightCount = nd.triCount - leftCount;
                if (leftCount == 0 || rightCount == 0 || taskCount == 256) break;
                const uint32_t n = next;
                next += 2;
                node[n] = {bLMin, bLMax, nd.leftFirst, leftCount};
                node[n + 1] = {bRMin, bRMax, j, rightCount};
                nd.leftFirst = n, nd.triCount = 0;
                if (depth < 5 && threaded) {
                    build(n, depth + 1);
                    build(n + 1, depth + 1);
                    break;
                }
                task[taskCount++] = n + 1, ni = n;
            }
            if (taskCount == 0) break;
            ni = task[--taskCount];
        }
    }
};

}  // namespace

extern "C" {

// gaussian_bvh.cu:150-155 + gaussian_bvh.cuh:20-26: three vec4 vertices per
// Gaussian (w = 0), rotations (w, x, y, z), activated scales.
int orc_bvh_fragments(const float* means, const float* rotations, const float* scales, uint64_t n, float* verts) {
    for (uint64_t i = 0; i < n; ++i) {
        const vec3<float> s{scales[3 * i], scales[3 * i + 1], scales[3 * i + 2]};
        const vec4<float> q{rotations[4 * i], rotations[4 * i + 1], rotations[4 * i + 2], rotations[4 * i + 3]};
        const mat3<float> cov = scale_rot_to_cov(s, q);
        const float ext[3] = {3.0f * std::sqrt(cov[0].x), 3.0f * std::sqrt(cov[1].y), 3.0f * std::sqrt(cov[2].z)};
        float lo[3], hi[3];
        for (int c = 0; c < 3; ++c) lo[c] = means[3 * i + c] - ext[c], hi[c] = means[3 * i + c] + ext[c];
        float* v = verts + 12 * i;
        const float tri[12] = {lo[0], lo[1], lo[2], 0.0f, hi[0], lo[1], lo[2], 0.0f, lo[0], hi[1], lo[2], 0.0f};
        std::memcpy(v, tri, sizeof(tri));
    }
    return 0;
}

static bool run_build(const float* verts, uint64_t n, Builder& b) {
    if (n == 0 || n > 0x7fffffffull) return false;
    b.frag.resize(n);
    b.prim.resize(n);
    b.node.resize(2 * n + 2);
    // PrepareBuild, tiny_bvh.h:2221-2290
    Node& root = b.node[0];
    root.leftFirst = 0, root.triCount = (uint32_t)n;
    root.mn = {{FAR, FAR, FAR}}, root.mx = {{-FAR, -FAR, -FAR}};
    for (uint64_t i = 0; i < n; ++i) {
        const float* t = verts + 12 * i;
        V3 v0{{t[0], t[1], t[2]}}, v1{{t[4], t[5], t[6]}}, v2{{t[8], t[9], t[10]}};
        b.frag[i].bmin = vmin3(v0, vmin3(v1, v2));
        b.frag[i].bmax = vmax3(v0, vmax3(v1, v2));
        root.mn = vmin3(root.mn, b.frag[i].bmin);
        root.mx = vmax3(root.mx, b.frag[i].bmax);
        b.prim[i] = (uint32_t)i;
    }
    for (int a = 0; a < 3; ++a) b.minDim.v[a] = (root.mx.v[a] - root.mn.v[a]) * 1e-20f;
    b.threaded = n >= MT_THRESHOLD;
    b.build(0, 0);
    return true;
}

// Leaves of the tinybvh build over `verts` (3 x vec4 per primitive), as
// primIdx position ranges: leaf_first[k], leaf_count[k] for the k-th leaf of
// a left-first walk.  Returns the number of leaves, -1 on bad input.
int64_t orc_bvh_leaves(const float* verts, uint64_t n, uint32_t* leaf_first, uint32_t* leaf_count) {
    Builder b;
    if (!run_build(verts, n, b)) return -1;
    int64_t k = 0;
    std::vector<uint32_t> st{0u};
    while (!st.empty()) {
        const uint32_t id = st.back();
        st.pop_back();
        const Node& nd = b.node[id];
        if (nd.triCount > 0) {
            leaf_first[k] = nd.leftFirst;
            leaf_count[k++] = nd.triCount;
        } else {
            st.push_back(nd.leftFirst + 1);
            st.push_back(nd.leftFirst);
        }
    }
    return k;
}

// perm[new] = old for the whole load-time reorder.
int orc_bvh_order(const float* means, const float* rotations, const float* scales, uint64_t n, uint32_t* perm) {
    if (n == 0) return 0;
    std::vector<float> verts(12 * n);
    orc_bvh_fragments(means, rotations, scales, n, verts.data());
    Builder b;
    if (!run_build(verts.data(), n, b)) return -1;
    // collect_leaf_order, gaussian_bvh.cu:14-34
    std::vector<uint32_t> leafOrder;
    leafOrder.reserve(n);
    std::vector<uint32_t> st{0u};
    while (!st.empty()) {
        const uint32_t id = st.back();
        st.pop_back();
        if (id >= b.node.size()) continue;
        const Node& nd = b.node[id];
        if (nd.triCount > 0) {
            if (nd.leftFirst >= n) continue;
            for (uint32_t i = 0; i < nd.triCount; ++i) leafOrder.push_back(nd.leftFirst + i);
        } else {
            st.push_back(nd.leftFirst);
            st.push_back(nd.leftFirst + 1);
        }
    }
    if (leafOrder.size() != n) return -1;
    // reorder, gaussian_bvh.cu:47-84: new[leafOrder[m]] = old[m]
    for (uint64_t m = 0; m < n; ++m) perm[leafOrder[m]] = (uint32_t)m;
    return 0;
}

}  // extern "C"
