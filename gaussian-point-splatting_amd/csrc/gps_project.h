// gps_project.h — per-Gaussian projection, EWA 2D covariance, screen bounds
// and SH colour in explicit scalar form for the gfx950 kernels.
//
// Follows src/core/rendering/passes/gaussian_point_splatting.cu:17-65
// (loadAndTransformGaussian), src/core/math/linalg.cuh:211-342 and
// src/core/math/sh.cuh:191-223.  Every sum is written in the association of
// glm 1.0.2's non-SIMD operators (mat3*vec3 = (c0*x + c1*y) + c2*z,
// mat4*vec4 = (c0*x + c1*y) + (c2*z + c3*w), mat3*mat3 columnwise likewise),
// with no FP contraction, so results are bit-identical to the CPU oracle.
#pragma once

#include "gps_detmath.h"

namespace gps {

// Camera constants of one frame (all column-major, glm layout).
struct FrameConsts {
    float V[16];        // view matrix
    float VP[16];       // proj * view
    float campos[3];    // affineInverse(view)[3]
    float tan_fov[2];
    float focal[2];
    float dilation;     // 0.3f * ss * ss
    float near_plane, far_plane;
    float K;            // points per kernel as float (ss^2 or 1)
    uint32_t point_cap; // (uint32_t)((W'*H') / (2*K))
    int32_t up_w, up_h;
    int32_t ss, spp, Ki;
    uint32_t frame_seed;
    int32_t occl, hier, cull_small, unbiased, reduce, only_prev;
    int32_t ss_shift;   // log2(ss) when ss is a power of two, else -1
    float vn2;          // >= |mat3(V)|_2^2 (footprint bounds of the occlusion tests)
    int32_t hiz_max;    // the Hi-Z chain is max-only (no min-reduced levels): the bounds apply
};

struct Projected {
    float view_z;
    float px, py, pz;   // NDC position (after divide)
    float pmx, pmy;     // pixel-space mean
    float ca, cb, cc;   // 2D covariance (dilated)
    float det;
    int32_t minx, miny, maxx, maxy;  // tile-aligned bounds
};

// linalg.cuh:319-342: R = mat3_cast(quat(w,x,y,z)), M = R*S, cov = M*M^T.
// Returns the full matrix in `C` (column-major 3x3).
//
// S is diagonal, so the reference's sums in M = R*S each hold one product and
// two products with an exact zero; they are left out here.  For a finite
// rotation (quaternion components below 2^63) a product with zero is +-0,
// and x + (+-0) is x for x != 0, so only the sign of a zero entry can differ.
// Signs of zeros never reach a result: cov enters the 2D covariance only
// through sums that end in "+ dilation" (> 0) or "+ 0.0f" (both zeros give
// +0), and the diagonal only through squares.  The same holds for the zero
// entries of the Jacobian in project_gaussian.
GPS_HD void scale_rot_to_cov(float sx, float sy, float sz, float qw, float qx, float qy, float qz, float C[9]) {
    const float xx = qx * qx, yy = qy * qy, zz = qz * qz;
    const float xz = qx * qz, xy = qx * qy, yz = qy * qz;
    const float wx = qw * qx, wy = qw * qy, wz = qw * qz;
    float R[9];
    R[0] = 1.0f - 2.0f * (yy + zz);
    R[1] = 2.0f * (xy + wz);
    R[2] = 2.0f * (xz - wy);
    R[3] = 2.0f * (xy - wz);
    R[4] = 1.0f - 2.0f * (xx + zz);
    R[5] = 2.0f * (yz + wx);
    R[6] = 2.0f * (xz + wy);
    R[7] = 2.0f * (yz - wx);
    R[8] = 1.0f - 2.0f * (xx + yy);
    const float s3[3] = {sx, sy, sz};
    float M[9];  // column j of R times s_j
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int r = 0; r < 3; ++r) M[3 * j + r] = R[3 * j + r] * s3[j];
    // M * transpose(M): column j of M^T is (M[0][j], M[1][j], M[2][j]).
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            C[3 * j + r] = (M[r] * M[j] + M[3 + r] * M[3 + j]) + M[6 + r] * M[6 + j];
}

// The centre part of loadAndTransformGaussian (gaussian_point_splatting.cu:
// 17-36): view-space position, near-plane test, NDC position and its test.
struct Center {
    float vx, vy, vz;  // view space
    float px, py, pz;  // NDC
};
GPS_HD bool project_center(const FrameConsts& f, float mx, float my, float mz, Center& c) {
    const float* V = f.V;
    // transformPoint4x3 (linalg.cuh:211-215)
    c.vx = ((V[0] * mx + V[4] * my) + V[8] * mz) + V[12];
    c.vy = ((V[1] * mx + V[5] * my) + V[9] * mz) + V[13];
    c.vz = ((V[2] * mx + V[6] * my) + V[10] * mz) + V[14];
    if (-c.vz <= 0.2f) return false;
    const float* P = f.VP;
    const float hx = (P[0] * mx + P[4] * my) + (P[8] * mz + P[12] * 1.0f);
    const float hy = (P[1] * mx + P[5] * my) + (P[9] * mz + P[13] * 1.0f);
    const float hz = (P[2] * mx + P[6] * my) + (P[10] * mz + P[14] * 1.0f);
    const float hw = (P[3] * mx + P[7] * my) + (P[11] * mz + P[15] * 1.0f);
    const float iw = 1.0f / hw;
    c.px = hx * iw;
    c.py = hy * iw;
    c.pz = hz * iw;
    return !(c.pz < -1.0f || c.pz > 1.0f);
}
// Pixel-space mean (upsampled pixels).
GPS_HD float pixel_mean(float p, int32_t size) { return (p * 0.5f + 0.5f) * (float)size; }

// The rest of loadAndTransformGaussian after project_center (its :37-65):
// 3D and 2D covariance, determinant test, pixel mean, tile bounds.
GPS_HD bool project_footprint(const FrameConsts& f, const Center& ce, float qw, float qx, float qy, float qz, float sx,
                              float sy, float sz, Projected& o) {
    const float* V = f.V;
    o.view_z = ce.vz;
    o.px = ce.px;
    o.py = ce.py;
    o.pz = ce.pz;
    const float vx = ce.vx, vy = ce.vy, vz = ce.vz;

    float C3[9];
    scale_rot_to_cov(sx, sy, sz, qw, qx, qy, qz, C3);

    // computeCov2D (linalg.cuh:221-249)
    const float limx = 1.3f * f.tan_fov[0];
    const float limy = 1.3f * f.tan_fov[1];
    const float txtz = vx / vz;
    const float tytz = vy / vz;
    const float tx = det_fmin(limx, det_fmax(-limx, txtz)) * vz;
    const float ty = det_fmin(limy, det_fmax(-limy, tytz)) * vz;
    // J = mat3(f0 / vz, 0, -(f0 tx) / vz^2,  0, f1 / vz, -(f1 ty) / vz^2,  0, 0, 0)
    const float j00 = f.focal[0] / vz, j02 = -(f.focal[0] * tx) / (vz * vz);
    const float j11 = f.focal[1] / vz, j12 = -(f.focal[1] * ty) / (vz * vz);
    // Wt = transpose(mat3(V)): Wt[c][r] = V[4r + c].  Columns 0 and 1 of
    // T = Wt * J without the products with J's zero entries (see
    // scale_rot_to_cov: only signs of zeros can differ).
    float T[6];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        T[r] = V[4 * r + 0] * j00 + V[4 * r + 2] * j02;
        T[3 + r] = V[4 * r + 1] * j11 + V[4 * r + 2] * j12;
    }
    // A = transpose(T) * C3, rows 0..1: A[j][r] = (T[r][0]*C[j].x + T[r][1]*C[j].y) + T[r][2]*C[j].z
    float A[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int r = 0; r < 2; ++r)
            A[3 * j + r] = (T[3 * r + 0] * C3[3 * j] + T[3 * r + 1] * C3[3 * j + 1]) + T[3 * r + 2] * C3[3 * j + 2];
    // cov = A * T: cov[j][r] = (A[0][r]*T[j].x + A[1][r]*T[j].y) + A[2][r]*T[j].z
    const float c00 = (A[0] * T[0] + A[3] * T[1]) + A[6] * T[2];
    const float c01 = (A[1] * T[0] + A[4] * T[1]) + A[7] * T[2];
    const float c11 = (A[1] * T[3] + A[4] * T[4]) + A[7] * T[5];

    o.ca = c00 + f.dilation;
    o.cb = c01 + 0.0f;
    o.cc = c11 + f.dilation;
    o.det = o.ca * o.cc - o.cb * o.cb;
    if (o.det <= 0.0f) return false;
    o.pmx = pixel_mean(o.px, f.up_w);
    o.pmy = pixel_mean(o.py, f.up_h);

    // compute2DGaussianBounds3DGS (linalg.cuh:251-296)
    // (float)(int) of the ceilings (the 3-sigma half extents, >= 0 or NaN);
    // ca and cc carry the dilation (>= 0.3), so the root needs no scaling of
    // tiny inputs
    const float ex = det_f2i_rz_float(det_ceil(3.0f * det_sqrt_normal(o.ca)));
    const float ey = det_f2i_rz_float(det_ceil(3.0f * det_sqrt_normal(o.cc)));
    const float mnx = o.pmx - ex, mny = o.pmy - ey;
    const float mxx = o.pmx + ex, mxy = o.pmy + ey;
    const int32_t tiles_x = (f.up_w + 15) / 16;
    const int32_t tiles_y = (f.up_h + 15) / 16;
    // the tile bounds only meet the tests and clamps below: converted after
    // clamping into [-1, tiles] / [0, tiles], with the same outcome as the
    // saturating conversion
    int32_t t0x = det_f2i_clamped(det_floor(mnx / 16.0f), -1.0f, (float)tiles_x);
    int32_t t0y = det_f2i_clamped(det_floor(mny / 16.0f), -1.0f, (float)tiles_y);
    int32_t t1x = det_f2i_clamped(det_ceil(mxx / 16.0f), 0.0f, (float)tiles_x);
    int32_t t1y = det_f2i_clamped(det_ceil(mxy / 16.0f), 0.0f, (float)tiles_y);
    if (t1x <= 0 || t1y <= 0 || t0x >= tiles_x || t0y >= tiles_y) return false;
    t0x = t0x < 0 ? 0 : t0x;
    t0y = t0y < 0 ? 0 : t0y;
    t1x = t1x > tiles_x ? tiles_x : t1x;
    t1y = t1y > tiles_y ? tiles_y : t1y;
    o.minx = t0x * 16;
    o.miny = t0y * 16;
    o.maxx = t1x * 16;
    o.maxy = t1y * 16;
    return true;
}

// loadAndTransformGaussian (gaussian_point_splatting.cu:17-65).  Returns false
// when the Gaussian is rejected.
GPS_HD bool project_gaussian(const FrameConsts& f, float mx, float my, float mz, float qw, float qx, float qy, float qz,
                             float sx, float sy, float sz, Projected& o) {
    Center ce;
    const bool front = project_center(f, mx, my, mz, ce);
    o.view_z = ce.vz;
    return front && project_footprint(f, ce, qw, qx, qy, qz, sx, sy, sz, o);
}

// SH basis constants (sh.cuh:137-169).
#define GPS_SH_C0 0.28209479177387814f
#define GPS_SH_C1 0.4886025119029199f
#define GPS_SH_C2_0 1.0925484305920792f
#define GPS_SH_C2_1 -1.0925484305920792f
#define GPS_SH_C2_2 0.31539156525252005f
#define GPS_SH_C2_3 -1.0925484305920792f
#define GPS_SH_C2_4 0.5462742152960396f
#define GPS_SH_C3_0 -0.5900435899266435f
#define GPS_SH_C3_1 2.890611442640554f
#define GPS_SH_C3_2 -0.4570457994644658f
#define GPS_SH_C3_3 0.3731763325901154f
#define GPS_SH_C3_4 -0.4570457994644658f
#define GPS_SH_C3_5 1.445305721320277f
#define GPS_SH_C3_6 -0.5900435899266435f

// computeColorFromSHFloat for one channel up to `degree`; `c(i)` returns
// coefficient i of the channel.  For coefficients above the scene's degree the
// reference adds exact zeros, which leave a non-zero running sum unchanged and
// only flip the sign of a zero one; both vanish after the final +0.5, so
// stopping at the scene's degree is result-identical for finite directions.
template <int DEG, typename Coef>
GPS_HD float sh_channel(float x, float y, float z, Coef c) {
    float result = GPS_SH_C0 * c(0);
    if (DEG > 0) {
        result = result - GPS_SH_C1 * y * c(1) + GPS_SH_C1 * z * c(2) - GPS_SH_C1 * x * c(3);
    }
    if (DEG > 1) {
        const float xx = x * x, yy = y * y, zz = z * z;
        const float xy = x * y, yz = y * z, xz = x * z;
        result = result + (GPS_SH_C2_0 * xy * c(4) + GPS_SH_C2_1 * yz * c(5) + GPS_SH_C2_2 * (2.0f * zz - xx - yy) * c(6) +
                           GPS_SH_C2_3 * xz * c(7) + GPS_SH_C2_4 * (xx - yy) * c(8));
        if (DEG > 2) {
            result = result + (GPS_SH_C3_0 * y * (3.0f * xx - yy) * c(9) + GPS_SH_C3_1 * xy * z * c(10) +
                               GPS_SH_C3_2 * y * (4.0f * zz - xx - yy) * c(11) +
                               GPS_SH_C3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * c(12) +
                               GPS_SH_C3_4 * x * (4.0f * zz - xx - yy) * c(13) + GPS_SH_C3_5 * z * (xx - yy) * c(14) +
                               GPS_SH_C3_6 * x * (xx - 3.0f * yy) * c(15));
        }
    }
    result = result + 0.5f;
    return (result < 0.0f) ? 0.0f : result;
}

// View direction normalize(pos - campos) (glm normalize = v * inversesqrt(dot(v, v))).
GPS_HD void view_dir(const FrameConsts& f, float mx, float my, float mz, float& x, float& y, float& z) {
    const float dx = mx - f.campos[0], dy = my - f.campos[1], dz = mz - f.campos[2];
    const float d = (dx * dx + dy * dy) + dz * dz;
    const float s = 1.0f / det_sqrt(d);
    x = dx * s;
    y = dy * s;
    z = dz * s;
}

}  // namespace gps
