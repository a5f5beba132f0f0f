// orc_ref.h — CPU restatement of the reference's device functions on the hot
// path, one function per reference function, in the reference's order of
// operations.  Fast-math intrinsics are replaced by the shared deterministic
// functions of gps_detmath.h (float) or libm (double).
//
// TEST INFRASTRUCTURE: part of the CPU oracle.
#pragma once

#include <atomic>

#include "../include/gps.h"
#include "orc_glm.h"

namespace orc {

// ------------------------------------------------------------ random.cuh
struct uint2_ {
    uint32_t x, y;
};
struct float2_ {
    float x, y;
};

// src/core/random/random.cuh:192-199
inline uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
// random.cuh:201-215
inline uint2_ makeSeed(uint32_t idx, uint32_t frame) {
    const uint32_t C1 = 0x9E3779B9u, C2 = 0xBB67AE85u, C3 = 0x3C6EF372u, C4 = 0xA54FF53Au;
    uint32_t h1 = hash32(idx ^ (frame * C1));
    uint32_t h2 = hash32((idx * C2) ^ frame);
    h1 ^= (h2 * C3);
    h2 ^= (h1 * C4);
    return {h1, h2};
}
// random.cuh:162-178
inline uint2_ pcg2d_uint(uint2_ s) {
    s.x = s.x * 1664525u + 1013904223u;
    s.y = s.y * 1664525u + 1013904223u;
    s.x += s.y * 1664525u;
    s.y += s.x * 1664525u;
    s.x ^= s.x >> 16u;
    s.y ^= s.y >> 16u;
    s.x += s.y * 1664525u;
    s.y += s.x * 1664525u;
    s.x ^= s.x >> 16u;
    s.y ^= s.y >> 16u;
    return s;
}
// random.cuh:181-184; RECIPROCAL_2e32 (2.3283064365387e-10f) is exactly 2^-32 in float.
inline float2_ pcg2d(uint2_& seed) {
    seed = pcg2d_uint(seed);
    const float R = 2.3283064365387e-10f;
    return {(float)seed.x * R, (float)seed.y * R};
}
// random.cuh:186-190
inline int stochastic_round(float x, float u) {
    int base = gps::det_f2i_rz(std::floor(x));
    float frac = x - static_cast<float>(base);
    int r = (u < frac) ? (int)((uint32_t)base + 1u) : base;
    return gmax(0, r);
}
static const float CUDART_PI_F = 3.141592654f;
// random.cuh:69-75
inline float2_ boxMuller(float u1, float u2) {
    float R = std::sqrt(gps::det_fmax(-2.0f * gps::det_logf(gclamp(u1, 1.e-37f, 1.0f)), 0.0f));
    float theta = 2.0f * CUDART_PI_F * u2;
    float s, c;
    gps::det_sincosf(theta, &s, &c);
    return {R * c, R * s};
}
// random.cuh:88-104
inline float dilog(float x) {
    float y = -6.09201442e-01f;
    y = std::fma(y, x, 1.79126616e+00f);
    y = std::fma(y, x, -2.14953223e+00f);
    y = std::fma(y, x, 1.26304372e+00f);
    y = std::fma(y, x, -4.59069895e-01f);
    y = std::fma(y, x, -1.87417414e-01f);
    y = std::fma(y, x, 1.99603130e+00f);
    y = std::fma(y, x, 5.99669467e-05f);
    float s = 1.0f - x;
    if (s > 0.0f) y += s * gps::det_logf(gps::det_fmax(s, 1e-37f));
    return y;
}
static const float LI2_MAX = 1.6449340668482264f;
// random.cuh:110-127
inline float inv_dilog(float x) {
    float t = gps::det_fmin(x / LI2_MAX, 1.0f);
    float y = -1.27463503e+01f;
    y = std::fma(y, t, 5.88993459e+01f);
    y = std::fma(y, t, -1.16025780e+02f);
    y = std::fma(y, t, 1.26945827e+02f);
    y = std::fma(y, t, -8.43108826e+01f);
    y = std::fma(y, t, 3.48799862e+01f);
    y = std::fma(y, t, -8.89606235e+00f);
    y = std::fma(y, t, 1.38640936e+00f);
    y = std::fma(y, t, -7.80640876e-01f);
    y = std::fma(y, t, 1.64841888e+00f);
    y = std::fma(y, t, -2.82836687e-05f);
    return y;
}
// random.cuh:132-141
inline float2_ correctedBoxMuller(float u1, float u2, float alpha) {
    float a = 1.0f / gps::det_fmax(alpha, 1.e-37f) * inv_dilog((1.0f - u1) * dilog(alpha));
    a = gclamp(a, 1.e-37f, 1.0f);
    float R = std::sqrt(gps::det_fmax(-2.0f * gps::det_logf(a), 0.0f));
    float theta = 2.0f * CUDART_PI_F * u2;
    float s, c;
    gps::det_sincosf(theta, &s, &c);
    return {R * c, R * s};
}
// poisson.cuh:12-38 (Giles QN3); __float2int_rn = round-half-even, saturating.
inline uint32_t poisson(uint2_& state, float lambda) {
    float2_ u = pcg2d(state);
    float w = boxMuller(u.x, u.y).x;
    float w2 = w * w;
    float w3 = w2 * w;
    float w4 = w2 * w2;
    float s = std::sqrt(lambda);
    float inv_s = 1.0f / s;
    float inv_l = 1.0f / lambda;
    float kf = lambda + s * w + (w2 - 1.0f) / 6.0f;
    kf += inv_s * (-(1.0f / 36.0f) * w - (1.0f / 72.0f) * w3);
    kf += inv_l * (-(8.0f / 405.0f) + (7.0f / 810.0f) * w2 + (1.0f / 270.0f) * w4);
    int ki = gps::det_f2i_rz(std::nearbyint(kf));
    return (uint32_t)gmax(ki, 0);
}

// ------------------------------------------------------------ linalg.cuh
// linalg.cuh:211-215
template <typename T> inline vec3<T> transformPoint4x3(vec3<T> p, const mat4<T>& m) {
    const mat3<T> B = mat3_from4(m);
    const vec3<T> o = {m[3][0], m[3][1], m[3][2]};
    return B * p + o;
}
// linalg.cuh:221-249 (EWA, Zwicker et al. 2002)
template <typename T>
inline vec3<T> computeCov2D(const mat4<T>& view, const mat3<T>& cov3D, vec3<T> mean, vec2<T> tan_fov, vec2<T> focal) {
    const mat3<T> W = mat3_from4(view);
    vec3<T> t = transformPoint4x3(mean, view);
    const T limx = T(1.3f) * tan_fov.x;
    const T limy = T(1.3f) * tan_fov.y;
    const T txtz = t.x / t.z;
    const T tytz = t.y / t.z;
    t.x = SF<T>::fmin(limx, SF<T>::fmax(-limx, txtz)) * t.z;
    t.y = SF<T>::fmin(limy, SF<T>::fmax(-limy, tytz)) * t.z;
    mat3<T> J;
    J.c[0] = {focal.x / t.z, T(0), -(focal.x * t.x) / (t.z * t.z)};
    J.c[1] = {T(0), focal.y / t.z, -(focal.y * t.y) / (t.z * t.z)};
    J.c[2] = {T(0), T(0), T(0)};
    const mat3<T> Tm = transpose(W) * J;
    const mat3<T> cov = transpose(Tm) * cov3D * Tm;
    return {cov[0][0], cov[0][1], cov[1][1]};
}
struct ivec2_ {
    int x, y;
};
// linalg.cuh:251-296; returns false when outside the tile grid.
template <typename T>
inline bool compute2DGaussianBounds3DGS(ivec2_& out_min, ivec2_& out_max, int res_x, int res_y, vec3<T> cov2d,
                                        vec2<T> pixel_mean) {
    const int TILE = 16;
    const T a = cov2d.x, c = cov2d.z;
    const int ex = gps::det_f2i_rz((float)SF<T>::ceil(T(3.0f) * SF<T>::sqrt(a)));
    const int ey = gps::det_f2i_rz((float)SF<T>::ceil(T(3.0f) * SF<T>::sqrt(c)));
    const vec2<T> mn = {pixel_mean.x - (T)ex, pixel_mean.y - (T)ey};
    const vec2<T> mx = {pixel_mean.x + (T)ex, pixel_mean.y + (T)ey};
    int tminx = gps::det_f2i_rz((float)SF<T>::floor(mn.x / T(TILE)));
    int tminy = gps::det_f2i_rz((float)SF<T>::floor(mn.y / T(TILE)));
    int tmaxx = gps::det_f2i_rz((float)SF<T>::ceil(mx.x / T(TILE)));
    int tmaxy = gps::det_f2i_rz((float)SF<T>::ceil(mx.y / T(TILE)));
    const int tiles_x = (res_x + TILE - 1) / TILE;
    const int tiles_y = (res_y + TILE - 1) / TILE;
    if (tmaxx <= 0 || tmaxy <= 0 || tminx >= tiles_x || tminy >= tiles_y) return false;
    // fmaxf/fminf on int operands: exact for the ranges that reach here.
    tminx = tminx < 0 ? 0 : tminx;
    tminy = tminy < 0 ? 0 : tminy;
    tmaxx = tmaxx > tiles_x ? tiles_x : tmaxx;
    tmaxy = tmaxy > tiles_y ? tiles_y : tmaxy;
    out_min = {tminx * TILE, tminy * TILE};
    out_max = {tmaxx * TILE, tmaxy * TILE};
    return true;
}
// linalg.cuh:301-307
inline vec3<float> conic(vec3<float> cov2d, float inv_det) {
    return {cov2d.z * inv_det, -cov2d.y * inv_det, cov2d.x * inv_det};
}
// linalg.cuh:309-317
inline float gaussian_value(vec3<float> conic, vec2<float> delta) {
    const float sigma = -0.5f * (conic.x * delta.x * delta.x + conic.z * delta.y * delta.y) - conic.y * delta.x * delta.y;
    if (sigma > 0.f) return 0.0f;
    return gps::det_expf(sigma);
}
// linalg.cuh:319-321: glm::toMat3(glm::quat{w=q.x, x=q.y, y=q.z, z=q.w})
// (glm/gtc/quaternion.inl mat3_cast)
template <typename T> inline mat3<T> quat_to_rotmat(vec4<T> q4) {
    const T w = q4.x, x = q4.y, y = q4.z, z = q4.w;
    const T qxx = x * x, qyy = y * y, qzz = z * z;
    const T qxz = x * z, qxy = x * y, qyz = y * z;
    const T qwx = w * x, qwy = w * y, qwz = w * z;
    mat3<T> R;
    R[0][0] = T(1) - T(2) * (qyy + qzz);
    R[0][1] = T(2) * (qxy + qwz);
    R[0][2] = T(2) * (qxz - qwy);
    R[1][0] = T(2) * (qxy - qwz);
    R[1][1] = T(1) - T(2) * (qxx + qzz);
    R[1][2] = T(2) * (qyz + qwx);
    R[2][0] = T(2) * (qxz + qwy);
    R[2][1] = T(2) * (qyz - qwx);
    R[2][2] = T(1) - T(2) * (qxx + qyy);
    return R;
}
// linalg.cuh:324-342: M = R * S; cov = M * transpose(M)
template <typename T> inline mat3<T> scale_rot_to_cov(vec3<T> s, vec4<T> q) {
    const mat3<T> R = quat_to_rotmat(q);
    mat3<T> S;
    S.c[0] = {s.x, T(0), T(0)};
    S.c[1] = {T(0), s.y, T(0)};
    S.c[2] = {T(0), T(0), s.z};
    const mat3<T> M = R * S;
    return M * transpose(M);
}
// linalg.cuh:406-411 (view z -> 28-bit depth); DEPTH_MAX is float(2^28-1) = 2^28.
inline uint64_t convert_view_depth_bits(float depth, float near_plane, float far_plane) {
    const float DEPTH_MAX_F = (float)((1ull << 28) - 1);
    const float t = gclamp((std::fabs(depth) - near_plane) / (far_plane - near_plane), 0.0f, 1.0f);
    return gps::det_f2u64_rz(t * DEPTH_MAX_F) & ((1ull << 28) - 1);
}
inline float convert_view_depth_float(uint64_t depth, float near_plane, float far_plane) {
    const float DEPTH_MAX_F = (float)((1ull << 28) - 1);
    return -((static_cast<float>(depth) / DEPTH_MAX_F) * (far_plane - near_plane) + near_plane);
}
// linalg.cuh:413-420
inline float view_to_projected_depth(float z_view, float near_plane, float far_plane) {
    float A = (far_plane + near_plane) / (far_plane - near_plane);
    float B = (2.0f * far_plane * near_plane) / (far_plane - near_plane);
    return A + B / z_view;
}

// ------------------------------------------------------------ sh.cuh
static const float SH_C0 = 0.28209479177387814f;
static const float SH_C1 = 0.4886025119029199f;
static const float SH_C2[] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f, -1.0925484305920792f,
                              0.5462742152960396f};
static const float SH_C3[] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f, 0.3731763325901154f,
                              -0.4570457994644658f, 1.445305721320277f, -0.5900435899266435f};
// sh.cuh:191-223, evaluated at full degree 3 (deg passed as 4 at
// gaussian_point_splatting.cu:269).
template <typename T> inline vec3<T> computeColorFromSHFloat(vec3<T> pos, vec3<T> campos, const T* shs) {
    vec3<T> dir = pos - campos;
    dir = normalize(dir);
    auto SH = [&](int i) { return vec3<T>{shs[3 * i], shs[3 * i + 1], shs[3 * i + 2]}; };
    vec3<T> result = T(SH_C0) * SH(0);
    const T x = dir.x, y = dir.y, z = dir.z;
    result = result - T(SH_C1) * y * SH(1) + T(SH_C1) * z * SH(2) - T(SH_C1) * x * SH(3);
    const T xx = x * x, yy = y * y, zz = z * z;
    const T xy = x * y, yz = y * z, xz = x * z;
    result = result + (T(SH_C2[0]) * xy * SH(4) + T(SH_C2[1]) * yz * SH(5) +
                       T(SH_C2[2]) * (T(2.0f) * zz - xx - yy) * SH(6) + T(SH_C2[3]) * xz * SH(7) +
                       T(SH_C2[4]) * (xx - yy) * SH(8));
    result = result + (T(SH_C3[0]) * y * (T(3.0f) * xx - yy) * SH(9) + T(SH_C3[1]) * xy * z * SH(10) +
                       T(SH_C3[2]) * y * (T(4.0f) * zz - xx - yy) * SH(11) +
                       T(SH_C3[3]) * z * (T(2.0f) * zz - T(3.0f) * xx - T(3.0f) * yy) * SH(12) +
                       T(SH_C3[4]) * x * (T(4.0f) * zz - xx - yy) * SH(13) + T(SH_C3[5]) * z * (xx - yy) * SH(14) +
                       T(SH_C3[6]) * x * (xx - T(3.0f) * yy) * SH(15));
    result = result + T(0.5f);
    return vmax(result, vec3<T>{T(0), T(0), T(0)});
}

// sh.cuh:55-96 (Sloan's real SH basis to degree 3)
template <typename T> inline void eval_sh_sloan_l3(T out_shs[16], const vec3<T> dir) {
    T x, y, z, z2, c0, s0, c1, s1, d, a, b;
    x = dir[0];
    y = dir[1];
    z = dir[2];
    z2 = z * z;
    d = T(0.282094792f);
    out_shs[0] = d;
    d = T(0.488602512f) * z;
    out_shs[2] = d;
    a = z2 - T(0.333333333f);
    d = T(0.946174696f) * a;
    out_shs[6] = d;
    b = z * (a - T(0.266666667f));
    d = T(1.865881663f) * b;
    out_shs[12] = d;
    c1 = x;
    s1 = y;
    d = T(-0.488602512f);
    out_shs[1] = s1 * d;
    out_shs[3] = c1 * d;
    d = T(-1.092548431f) * z;
    out_shs[5] = s1 * d;
    out_shs[7] = c1 * d;
    a = z2 - T(0.2f);
    d = T(-2.285228997f) * a;
    out_shs[11] = s1 * d;
    out_shs[13] = c1 * d;
    c0 = x * c1 - y * s1;
    s0 = y * c1 + x * s1;
    d = T(0.546274215f);
    out_shs[4] = s0 * d;
    out_shs[8] = c0 * d;
    d = T(1.445305721f) * z;
    out_shs[10] = s0 * d;
    out_shs[14] = c0 * d;
    c1 = x * c0 - y * s0;
    s1 = y * c0 + x * s0;
    d = T(-0.59004359f);
    out_shs[9] = s1 * d;
    out_shs[15] = c1 * d;
}
// sh.cuh:257-273, with the 16 coefficients already decoded by CompactSH::get
// (shs[3 * i + c] = get(i)[c]).
template <typename T> inline vec3<T> computeColorFromCompactSH(vec3<T> pos, vec3<T> campos, const T* shs) {
    vec3<T> dir = pos - campos;
    dir = normalize(dir);
    T eval_sh[16];
    eval_sh_sloan_l3(eval_sh, dir);
    vec3<T> result = {T(0), T(0), T(0)};
    for (int i = 0; i < 16; ++i) {
        const vec3<T> c = {shs[3 * i], shs[3 * i + 1], shs[3 * i + 2]};
        result = result + eval_sh[i] * c;
    }
    result = result + T(0.5f);
    return vmax(result, vec3<T>{T(0), T(0), T(0)});
}

// ------------------------------------------ gaussian_types.cuh, compressed scene
// glm::mix(x, y, a) = x * (1 - a) + y * a (func_common.inl:111-162)
inline float mix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
inline vec3<float> mix(vec3<float> x, vec3<float> y, vec3<float> a) {
    return {mix(x.x, y.x, a.x), mix(x.y, y.y, a.y), mix(x.z, y.z, a.z)};
}
// :212-235, host path (1 / sqrt, sin, cos)
inline vec4<float> quat_iem(const vec3<float> v) {
    const float d = v.x * v.x + v.y * v.y + v.z * v.z;
    const float t = 1.0f / gps::det_sqrt(d + 1e-8f);
    const float a = 1.57079632679f * t * d;
    float s, c;
    gps::det_sincosf(a, &s, &c);
    const float k = s * t;
    return {k * v.x, k * v.y, k * v.z, c};
}
struct GaussianGeometry {
    vec3<float> mean;
    vec4<float> quat;
    vec3<float> scale;
};
// :323-374 (__expf replaced by the deterministic exp)
inline GaussianGeometry decode_gaussian_geometry(const uint32_t qg[4], const gps_quantized_bounds& scene) {
    GaussianGeometry result{};
    const uint32_t px = qg[0] & 0x7FFFFFu;
    const uint32_t py_low9 = (qg[0] >> 23) & 0x1FFu;
    const uint32_t py_high13 = qg[1] & 0x1FFFu;
    const uint32_t pz_low19 = (qg[1] >> 13) & 0x7FFFFu;
    const uint32_t pz_high4 = qg[2] & 0xFu;
    const uint32_t py = py_low9 | (py_high13 << 9);
    const uint32_t pz = pz_low19 | (pz_high4 << 19);
    const float fx = float(px) / float((1u << 23) - 1u);
    const float fy = float(py) / float((1u << 22) - 1u);
    const float fz = float(pz) / float((1u << 23) - 1u);
    const vec3<float> minPos = {scene.min_pos[0], scene.min_pos[1], scene.min_pos[2]};
    const vec3<float> maxPos = {scene.max_pos[0], scene.max_pos[1], scene.max_pos[2]};
    result.mean = mix(minPos, maxPos, vec3<float>{fx, fy, fz});
    const uint32_t rx = (qg[2] >> 4) & 0x3FFu;
    const uint32_t ry = (qg[2] >> 14) & 0x3FFu;
    const uint32_t rz_low8 = (qg[2] >> 24) & 0xFFu;
    const uint32_t rz_high2 = qg[3] & 0x3u;
    const uint32_t rz = rz_low8 | (rz_high2 << 8);
    const vec3<float> vr = vec3<float>{float(rx), float(ry), float(rz)} / 1023.0f;
    const vec3<float> v = vr * 2.0f - 1.0f;
    result.quat = quat_iem(v);
    const uint32_t sx = (qg[3] >> 2) & 0x3FFu;
    const uint32_t sy = (qg[3] >> 12) & 0x3FFu;
    const uint32_t sz = (qg[3] >> 22) & 0x3FFu;
    const vec3<float> st = vec3<float>{float(sx), float(sy), float(sz)} / 1023.0f;
    const vec3<float> minScale = {scene.min_log_scale[0], scene.min_log_scale[1], scene.min_log_scale[2]};
    const vec3<float> maxScale = {scene.max_log_scale[0], scene.max_log_scale[1], scene.max_log_scale[2]};
    result.scale = mix(minScale, maxScale, st);
    result.scale = {gps::det_expf(result.scale.x), gps::det_expf(result.scale.y), gps::det_expf(result.scale.z)};
    return result;
}
// :188-191 (normalized_byte_to_float, the decode_opacity of a byte)
inline float normalized_byte_to_float(const uint8_t value) { return static_cast<float>(value) / 255.0f; }
// CompactSH::get, sh.cuh:41-47
inline vec3<float> compact_sh_get(int idx, const float base[3], const uint8_t sh[45],
                                  const gps_quantized_bounds& scene) {
    if (idx <= 0) return {base[0], base[1], base[2]};
    --idx;
    constexpr float r = 1.0f / 255.0f;
    const vec3<float> vec = {sh[3 * idx] * r, sh[3 * idx + 1] * r, sh[3 * idx + 2] * r};
    const vec3<float> mn = {scene.min_sh[3 * idx], scene.min_sh[3 * idx + 1], scene.min_sh[3 * idx + 2]};
    const vec3<float> mx = {scene.max_sh[3 * idx], scene.max_sh[3 * idx + 1], scene.max_sh[3 * idx + 2]};
    return mix(mn, mx, vec);
}

// ------------------------------------------------- post_processing_kernels.cuh
// :71-80
inline uint32_t mortonEncode(uint32_t x, uint32_t y) {
    auto spread = [](uint32_t a) {
        a = (a | (a << 8)) & 0x00FF00FF;
        a = (a | (a << 4)) & 0x0F0F0F0F;
        a = (a | (a << 2)) & 0x33333333;
        a = (a | (a << 1)) & 0x55555555;
        return a;
    };
    return spread(x) | (spread(y) << 1);
}
// :82-104 (MORTON_ORDER_IMAGE_BUFFER, 64x64 tiles)
inline int posToMemory(int x, int y, int p, int numPasses, int width, int height) {
    const int tx = x >> 6, ty = y >> 6;
    const int lx = x & 63, ly = y & 63;
    const uint32_t code = mortonEncode(lx, ly);
    const int tiles_per_row = (width + 63) >> 6;
    (void)height;
    const int tile_index = ty * tiles_per_row + tx;
    return (tile_index * 64 * 64 + (int)code) * numPasses + p;
}

// gaussian_types.cuh:638-645 (colour cache -> 12-bit packed)
inline uint64_t get_color_packed(vec3<float> c) {
    uint32_t r12 = gps::det_f2u_rz(c.x * 255.0f) & 0xFFF;
    uint32_t g12 = gps::det_f2u_rz(c.y * 255.0f) & 0xFFF;
    uint32_t b12 = gps::det_f2u_rz(c.z * 255.0f) & 0xFFF;
    return (uint64_t(r12) << 24) | (uint64_t(g12) << 12) | (uint64_t(b12) << 0);
}
// config.h:36-41
constexpr uint64_t COLOR_MASK = (1ull << 36) - 1;
constexpr uint64_t DEPTH_MASK = ((1ull << 28) - 1) << 36;

// Events of the biased splat whose outcome depends on the order in which
// samples reach a word (see splat_sample_biased).
struct BiasedEvents {
    std::atomic<uint64_t> zero_colour{0};  // a sample of packed colour 0 passed the pre-check
    std::atomic<uint64_t> wrapped{0};      // an equal-depth add carried out of bit 35
};

// post_processing_kernels.cuh:13-42, atomicAddColor, on any word type with
// load() / compare_exchange_strong() (std::atomic_ref<uint64_t>): a nearer
// sample replaces the word, a farther one is ignored, one at equal depth adds
// its 36 colour bits to the stored ones modulo 2^36 (a 12-bit channel that
// overflows carries into the next; the top one is dropped).  Returns false
// only when add_if_equal is false and the depths are equal.
template <typename Word> inline bool atomicAddColor(Word& w, uint64_t value, bool add_if_equal,
                                                   BiasedEvents* ev = nullptr) {
    const uint64_t target_depth = value & DEPTH_MASK;
    uint64_t old = w.load(std::memory_order_relaxed), assumed;
    do {
        assumed = old;
        const uint64_t assumed_depth = assumed & DEPTH_MASK;
        uint64_t new_value = value;
        if (target_depth > assumed_depth) {
            return true;
        } else if (target_depth == assumed_depth) {
            if (!add_if_equal) return false;
            const uint64_t new_lower = (assumed & COLOR_MASK) + (value & COLOR_MASK);
            if (ev && (new_lower >> 36)) ev->wrapped.fetch_add(1, std::memory_order_relaxed);
            new_value = (assumed & DEPTH_MASK) | (new_lower & COLOR_MASK);
        } else {
            new_value = value;
        }
        old = assumed;
        w.compare_exchange_strong(old, new_value, std::memory_order_relaxed);
    } while (assumed != old);
    return true;
}

// gaussian_point_splatting.cu:130-138, biased branch: the racy pre-check
// `imageBuffer[index] <= view_depth_bits` (a stored word that is nearer, or
// at the same depth with colour 0, skips the sample), then atomicAddColor
// with add_if_equal.  The word's final value is independent of the order of
// its samples unless a sample of colour 0 arrives first at the word's final
// depth (the pre-check then skips every later add there) or the colours
// summed at one depth wrap to 0 modulo 2^36; `ev` counts the samples that
// could start either case.
template <typename Word> inline void splat_sample_biased(Word& w, uint64_t value, BiasedEvents* ev = nullptr) {
    const uint64_t view_depth_bits = value & DEPTH_MASK;
    if (w.load(std::memory_order_relaxed) <= view_depth_bits) return;
    if (ev && (value & COLOR_MASK) == 0) ev->zero_colour.fetch_add(1, std::memory_order_relaxed);
    atomicAddColor(w, value, true, ev);
}

// sh.cuh:124-133
inline uint64_t vec3To36Bits(vec3<float> rgb) {
    const float a = 255.f;
    return ((gps::det_f2u64_rz(rgb.x * a) & 0xFFF) << 24) | ((gps::det_f2u64_rz(rgb.y * a) & 0xFFF) << 12) |
           ((gps::det_f2u64_rz(rgb.z * a) & 0xFFF));
}

}  // namespace orc
