// gps_hiz.h — the Hi-Z occlusion tests, host and device: the per-Gaussian
// test of the preprocess and the conservative group test that lets the
// preprocess skip whole waves.  Plain C++ with no HIP dependency, so the
// soundness of the group test can be checked on the CPU against the same
// per-Gaussian code the GPU runs.
#pragma once

#include <math.h>
#include <stdint.h>

#include "gps_project.h"

namespace gps {

// Hi-Z pyramid (coarsest level 0 first), src/core/rendering/occlusion/depth_mip_chain.cuh:10-71.
#define GPS_HIZ_MAX_LEVELS 16
struct HiZView {
    float* mip;
    int32_t width, height, levels;
    int32_t lw[GPS_HIZ_MAX_LEVELS];
    int32_t lh[GPS_HIZ_MAX_LEVELS];
    uint32_t off[GPS_HIZ_MAX_LEVELS];
};

// Reciprocal and square root for the bounds only (never for a rendered
// value): the hardware approximations on the GPU (within 1 ulp), the correctly
// rounded ones on the host; the bounds' 1e-5 relative margins cover either.
GPS_HD float hz_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
GPS_HD float hz_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

GPS_HD int hz_imin(int a, int b) { return a < b ? a : b; }
GPS_HD int hz_imax(int a, int b) { return a > b ? a : b; }

// ceil(log2(d / 2)) for integer d >= 1, exactly as the float expression of
// depth_mip_chain.cuh:186 evaluates it (log2 of a power of two is exact; any
// other d < 2^24 is far from an integer).
GPS_HD int ceil_log2_half(int d) {
    if (d <= 1) return -1;
    const int cl = 32 - __builtin_clz((uint32_t)(d - 1));  // ceil(log2(d))
    return cl - 1;
}

// depth_mip_chain.cuh:172-215
GPS_HD bool hiz_occluded(const HiZView& h, float depth, int minx, int miny, int maxx, int maxy) {
    minx = hz_imax(0, minx);
    miny = hz_imax(0, miny);
    maxx = hz_imin(h.width - 1, maxx);
    maxy = hz_imin(h.height - 1, maxy);
    const int w = maxx - minx + 1;
    const int hh = maxy - miny + 1;
    const int d = hz_imax(w, hh);
    int lod = ceil_log2_half(d);
    lod = hz_imin(hz_imax(lod, 0), h.levels - 1);
    const int l = (h.levels - 1) - lod;
    const int wl = h.lw[l], hl = h.lh[l];
    // The tile is max(1, min(size, 2^k)) texels with k = levels - l - 1 >= 0:
    // a power of two (a shift), or the whole image, which the coordinates
    // (clamped to [0, size - 1] above) divide to 0.
    const int k = h.levels - l - 1;
    const bool kx = (1 << k) <= h.width, ky = (1 << k) <= h.height;
    const int tx0 = hz_imax(0, hz_imin(wl - 1, kx ? minx >> k : 0));
    const int ty0 = hz_imax(0, hz_imin(hl - 1, ky ? miny >> k : 0)) * wl;
    const int tx1 = hz_imax(0, hz_imin(wl - 1, kx ? maxx >> k : 0));
    const int ty1 = hz_imax(0, hz_imin(hl - 1, ky ? maxy >> k : 0)) * wl;
    const float* m = h.mip + h.off[l];
    const float zmax = det_fmax(m[tx0 + ty0], det_fmax(m[tx1 + ty0], det_fmax(m[tx0 + ty1], m[tx1 + ty1])));
    return zmax < depth;
}

// A supersampled pixel bound in Hi-Z pixels: x / ss, as a shift when ss is a
// power of two.  For negative x the shift rounds down where the division
// truncates; hiz_occluded clamps both to 0, and only the minima can be
// negative (a Gaussian with a negative maximum was rejected before).
GPS_HD int div_ss(const FrameConsts& f, int x) { return f.ss_shift >= 0 ? x >> f.ss_shift : x / f.ss; }

// Bounds of a group of GPS_GROUP consecutive Gaussians (one preprocess wave),
// from the attributes exactly as the projection decodes them.
#define GPS_GROUP 64
struct GroupBox {
    float lo[3];   // box of the centres
    float lam;     // >= largest eigenvalue of every member's 3D covariance
    float hi[3];
    uint32_t ok;   // 0: a member has a non-finite centre, rotation or scale (never culled)
    float sig[3];  // >= sqrt(cov3D[i][i]) of every member, i = x, y, z
    float pad;
};
static_assert(sizeof(GroupBox) == 48, "three dwordx4 per group");

// Group bounds of one member (combined over the group by min / max / and);
// a = centre xyz, quaternion wxyz, scale xyz as the projection decodes them.
// mat3_cast of a quaternion of squared norm n is (1 - n) I + n R_unit, so
// |R|_2 <= |1 - n| + n and |R S|_2 <= (|1 - n| + n) max|s|; the diagonal of
// the covariance is the one scale_rot_to_cov computes.
GPS_HD bool group_member_bounds(const float a[10], float* lam, float sig[3]) {
    const float qw = a[3], qx = a[4], qy = a[5], qz = a[6], sx = a[7], sy = a[8], sz = a[9];
    float C[9];
    scale_rot_to_cov(sx, sy, sz, qw, qx, qy, qz, C);
    const double n = (((double)qw * qw + (double)qx * qx) + (double)qy * qy) + (double)qz * qz;
    const double rs = (fabs(1.0 - n) + n) * fmax(fabs((double)sx), fmax(fabs((double)sy), fabs((double)sz)));
    *lam = (float)(rs * rs * (1.0 + 1e-6));
    for (int i = 0; i < 3; ++i) sig[i] = (float)(sqrt(fmax(0.0, (double)C[4 * i])) * (1.0 + 1e-6));
    bool ok = isfinite(*lam) && isfinite(sig[0]) && isfinite(sig[1]) && isfinite(sig[2]);
    for (int i = 0; i < 10; ++i) ok = ok && isfinite(a[i]);
    return ok;
}

// Upper bound on |mat3(V)|_2^2 for group_occluded:
// lambda_max(A^T A) <= 1 + |A^T A - I|_F (A = the view's rotation part).
GPS_HD double view_norm2_bound(const float V[16]) {
    double g = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double d = 0.0;
            for (int k = 0; k < 3; ++k) d += (double)V[4 * i + k] * (double)V[4 * j + k];
            if (i == j) d -= 1.0;
            g += d * d;
        }
    return (1.0 + sqrt(g)) * (1.0 + 1e-9);
}

// Hi-Z texels covering [X0, X1] x [Y0, Y1] (Hi-Z pixels, inside the image)
// at the level of lod >= the lod a rectangle inside it would use: true when
// their maximum is below `depth`.  Reads at most 2 x 2 texels: a span of d <=
// 2^(lod+1) pixels covers at most 3 texels of 2^lod and 2 of 2^(lod+1).
GPS_HD bool hiz_rect_below(const HiZView& h, int X0, int Y0, int X1, int Y1, float depth) {
    int lod = ceil_log2_half(hz_imax(X1 - X0 + 1, Y1 - Y0 + 1));
    lod = hz_imin(hz_imax(lod, 0), h.levels - 1);
    if (((X1 >> lod) - (X0 >> lod) > 1 || (Y1 >> lod) - (Y0 >> lod) > 1) && lod < h.levels - 1) ++lod;
    const int l = (h.levels - 1) - lod;
    const int lw = h.lw[l], lh = h.lh[l];
    const int tx0 = hz_imin(lw - 1, X0 >> lod), tx1 = hz_imin(lw - 1, X1 >> lod);
    const int ty0 = hz_imin(lh - 1, Y0 >> lod), ty1 = hz_imin(lh - 1, Y1 >> lod);
    if (tx1 - tx0 > 1 || ty1 - ty0 > 1) return false;
    const float* m = h.mip + h.off[l];
    const float a = m[ty0 * lw + tx0], b = m[ty0 * lw + tx1], c = m[ty1 * lw + tx0], d = m[ty1 * lw + tx1];
    // a NaN texel fails every comparison: not occluded
    return a < depth && b < depth && c < depth && d < depth;
}

// The conservative group test.  True only when every member of the group
// that reaches the per-Gaussian Hi-Z test of the preprocess
// (gaussian_point_splatting.cu:226-242; hiz_occluded above, after the
// projection, bounds and small-eigenvalue steps) is rejected by it.  Why:
//  1. A member's test reads the four corner texels of its footprint rectangle
//     (tile-aligned 3-sigma bounds of the dilated 2D covariance, in Hi-Z
//     pixels) at lod_m = ceil(log2(longest side / 2)) and rejects it when
//     their maximum is below its centre's NDC depth.
//  2. That rectangle lies inside the group rectangle R below: the centres'
//     screen positions lie between the box corners' (a projection is
//     linear-fractional, so over a box in front of the camera its extremes
//     are at corners), a member's half-width is at most 3 sqrt(ca) + 1 with
//     ca bounded through the Jacobian (focal / depth over the box, the x/z
//     term clamped as the projection clamps it) and the group's covariance
//     bounds: ca <= min((sum_r |T_r| sig_r)^2, |V|^2 |J|^2 lam) + dilation,
//     T = transpose(mat3(V)) J; and the tile snapping is monotone.
//  3. lod_m <= lod_R, and the chain is a max pyramid (each texel the maximum
//     of its 2x2 children, ceil(size / 2) texels per level), so every texel a
//     member reads is at most the texel at lod_R (or one level up) covering
//     it, one of the at most 2 x 2 texels hiz_rect_below reads.
//  4. A member's NDC depth is at least the smallest over the box corners.
// Rounding: the corners go through the members' own float operations, so
// each computed corner value and each member value is within the same bound
// of its exact value (a few ulps of the magnitude of its terms, the
// magnitudes taken over the box); every comparison is widened by twice that.
// `vn2` >= |mat3(V)|_2^2.  Needs a max-only chain (no min-reduced levels).
GPS_HD bool group_occluded(const GroupBox& bx, const FrameConsts& f, const HiZView& h) {
    if (!bx.ok) return false;
    const float u = 5.9604645e-8f;  // 2^-24, unit roundoff of float
    const float* V = f.V;
    const float* P = f.VP;
    const float ax = fmaxf(fabsf(bx.lo[0]), fabsf(bx.hi[0]));
    const float ay = fmaxf(fabsf(bx.lo[1]), fabsf(bx.hi[1]));
    const float az = fmaxf(fabsf(bx.lo[2]), fabsf(bx.hi[2]));
    // magnitude of the terms of row r of a 4x4 float matrix applied to a point of the box
    auto mag = [&](const float* M, int r) {
        return ((fabsf(M[r]) * ax + fabsf(M[4 + r]) * ay) + fabsf(M[8 + r]) * az) + fabsf(M[12 + r]);
    };
    const float sx = mag(P, 0), sy = mag(P, 1), sz = mag(P, 2), sw = mag(P, 3);
    const float svx = mag(V, 0), svy = mag(V, 1), svz = mag(V, 2);
    float pxlo = 3.0e38f, pxhi = -3.0e38f, pylo = 3.0e38f, pyhi = -3.0e38f, zlo = 3.0e38f, wlo = 3.0e38f, dlo = 3.0e38f;
    float pxa = 0.0f, pya = 0.0f, txa = 0.0f, tya = 0.0f;
    for (int i = 0; i < 8; ++i) {
        const float cx = (i & 1) ? bx.hi[0] : bx.lo[0];
        const float cy = (i & 2) ? bx.hi[1] : bx.lo[1];
        const float cz = (i & 4) ? bx.hi[2] : bx.lo[2];
        // the operations of project_center
        const float vx = ((V[0] * cx + V[4] * cy) + V[8] * cz) + V[12];
        const float vy = ((V[1] * cx + V[5] * cy) + V[9] * cz) + V[13];
        const float vz = ((V[2] * cx + V[6] * cy) + V[10] * cz) + V[14];
        const float hx = (P[0] * cx + P[4] * cy) + (P[8] * cz + P[12] * 1.0f);
        const float hy = (P[1] * cx + P[5] * cy) + (P[9] * cz + P[13] * 1.0f);
        const float hz = (P[2] * cx + P[6] * cy) + (P[10] * cz + P[14] * 1.0f);
        const float hw = (P[3] * cx + P[7] * cy) + (P[11] * cz + P[15] * 1.0f);
        if (!(hw > 0.0f) || !(vz < 0.0f)) return false;  // the box reaches behind the camera
        const float iw = 1.0f / hw, ivz = hz_rcp(-vz);
        const float px = hx * iw, py = hy * iw, pz = hz * iw;
        pxlo = fminf(pxlo, px);
        pxhi = fmaxf(pxhi, px);
        pylo = fminf(pylo, py);
        pyhi = fmaxf(pyhi, py);
        pxa = fmaxf(pxa, fabsf(px));
        pya = fmaxf(pya, fabsf(py));
        zlo = fminf(zlo, pz);
        wlo = fminf(wlo, hw);
        dlo = fminf(dlo, -vz);
        txa = fmaxf(txa, fabsf(vx) * ivz);
        tya = fmaxf(tya, fabsf(vy) * ivz);
    }
    const float wl = wlo - 16.0f * u * sw;  // every member's float clip w is at least this
    if (!(wl > 0.0f)) return false;
    const float dl = fmaxf(dlo - 16.0f * u * svz, 0.2f);  // members reaching the test have -vz > 0.2
    const float iwl = hz_rcp(wl) * 1.00001f, idl = hz_rcp(dl) * 1.00001f;
    const float ez = 16.0f * u + 32.0f * u * (sz + sw) * iwl;  // twice a value's NDC depth error (|pz| <= 1 there)
    const float epx = (float)f.up_w * (16.0f * u * (pxa + 1.0f) + 32.0f * u * (sx + pxa * sw) * iwl) + 1.0f;
    const float epy = (float)f.up_h * (16.0f * u * (pya + 1.0f) + 32.0f * u * (sy + pya * sw) * iwl) + 1.0f;
    // Jacobian: |J00| = f0 / |vz|, |J02| = f0 |clamp(vx / vz)| / |vz| (and the y row likewise)
    const float tx = fminf(1.3f * f.tan_fov[0], txa * 1.00001f + 32.0f * u * (svx + txa * svz) * idl) * 1.00001f;
    const float ty = fminf(1.3f * f.tan_fov[1], tya * 1.00001f + 32.0f * u * (svy + tya * svz) * idl) * 1.00001f;
    const float j0 = f.focal[0] * idl * 1.00001f, j2 = j0 * tx;
    const float j4 = f.focal[1] * idl * 1.00001f, j5 = j4 * ty;
    // T[r] = V[4r] J00 + V[4r+2] J02, T[3+r] = V[4r+1] J11 + V[4r+2] J12
    float a_sig = 0.0f, c_sig = 0.0f;
    for (int r = 0; r < 3; ++r) {
        a_sig += (fabsf(V[4 * r]) * j0 + fabsf(V[4 * r + 2]) * j2) * bx.sig[r];
        c_sig += (fabsf(V[4 * r + 1]) * j4 + fabsf(V[4 * r + 2]) * j5) * bx.sig[r];
    }
    const float a_lam = f.vn2 * (j0 * j0 + j2 * j2) * bx.lam;
    const float c_lam = f.vn2 * (j4 * j4 + j5 * j5) * bx.lam;
    const float ca = fminf(a_sig * a_sig, a_lam) * 1.0001f + f.dilation * 1.00001f;
    const float cc = fminf(c_sig * c_sig, c_lam) * 1.0001f + f.dilation * 1.00001f;
    const float ex = 3.0f * hz_sqrt(ca) * 1.00001f + 1.0f;
    const float ey = 3.0f * hz_sqrt(cc) * 1.00001f + 1.0f;
    const float mxlo = pixel_mean(pxlo, f.up_w) - epx - ex, mxhi = pixel_mean(pxhi, f.up_w) + epx + ex;
    const float mylo = pixel_mean(pylo, f.up_h) - epy - ey, myhi = pixel_mean(pyhi, f.up_h) + epy + ey;
    if (!(fabsf(mxlo) < 1e6f && fabsf(mxhi) < 1e6f && fabsf(mylo) < 1e6f && fabsf(myhi) < 1e6f)) return false;
    // member bounds: 16 floor(fl(pm - e) / 16) .. 16 ceil(fl(pm + e) / 16),
    // tiles clamped to the screen; floor and ceil are monotone
    const float xmax = 16.0f * (float)((f.up_w + 15) / 16), ymax = 16.0f * (float)((f.up_h + 15) / 16);
    const float x0 = fminf(xmax, fmaxf(0.0f, 16.0f * floorf(mxlo * 0.0625f)));
    const float x1 = fminf(xmax, fmaxf(0.0f, 16.0f * ceilf(mxhi * 0.0625f)));
    const float y0 = fminf(ymax, fmaxf(0.0f, 16.0f * floorf(mylo * 0.0625f)));
    const float y1 = fminf(ymax, fmaxf(0.0f, 16.0f * ceilf(myhi * 0.0625f)));
    const int X0 = hz_imin(h.width - 1, div_ss(f, (int)x0)), X1 = hz_imin(h.width - 1, div_ss(f, (int)x1));
    const int Y0 = hz_imin(h.height - 1, div_ss(f, (int)y0)), Y1 = hz_imin(h.height - 1, div_ss(f, (int)y1));
    // every member's pz >= zlo - ez; one more ulp for the subtraction
    const float d = (zlo - ez) - 2.0f * u * fabsf(zlo);
    return hiz_rect_below(h, X0, Y0, X1, Y1, d);
}

}  // namespace gps
