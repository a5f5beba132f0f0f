// orc_glm.h — the few glm 1.0.2 value types and operators the reference's hot
// path uses, restated with glm's non-SIMD evaluation order so float results
// follow a fixed association (glm/detail/type_mat3x3.inl mul3x3, type_mat4x4.inl
// operator*(mat4, vec4) and mul4x4, func_geometric.inl dot/normalize,
// func_matrix.inl inv3x3 / compute_inverse<4,4>, gtc/matrix_inverse.inl
// affineInverse).  Templated on the scalar so the oracle's projection / SH
// path can run in FP32 (bit contract with the GPU) and FP64 (accuracy
// reference).
//
// TEST INFRASTRUCTURE: part of the CPU oracle; only tests/, bench.py's CPU
// baseline and the smoke check use it.
#pragma once

#include <cmath>
#include <cstdint>

#include "../gaussian-point-splatting_amd/csrc/gps_detmath.h"

namespace orc {

// Scalar functions: float -> the shared deterministic math, double -> libm.
template <typename T> struct SF;
template <> struct SF<float> {
    static float sqrt(float x) { return gps::det_sqrt(x); }
    static float exp(float x) { return gps::det_expf(x); }
    static float log(float x) { return gps::det_logf(x); }
    static void sincos(float x, float* s, float* c) { gps::det_sincosf(x, s, c); }
    static float fmin(float a, float b) { return gps::det_fmin(a, b); }
    static float fmax(float a, float b) { return gps::det_fmax(a, b); }
    static float floor(float x) { return gps::det_floor(x); }
    static float ceil(float x) { return gps::det_ceil(x); }
};
template <> struct SF<double> {
    static double sqrt(double x) { return std::sqrt(x); }
    static double exp(double x) { return std::exp(x); }
    static double log(double x) { return std::log(x); }
    static void sincos(double x, double* s, double* c) { *s = std::sin(x); *c = std::cos(x); }
    static double fmin(double a, double b) { return std::fmin(a, b); }
    static double fmax(double a, double b) { return std::fmax(a, b); }
    static double floor(double x) { return std::floor(x); }
    static double ceil(double x) { return std::ceil(x); }
};

template <typename T> inline T gmin(T x, T y) { return (y < x) ? y : x; }
template <typename T> inline T gmax(T x, T y) { return (x < y) ? y : x; }
template <typename T> inline T gclamp(T x, T lo, T hi) { return gmin(gmax(x, lo), hi); }

template <typename T> struct vec2 {
    T x, y;
};
template <typename T> struct vec3 {
    T x, y, z;
    T& operator[](int i) { return (&x)[i]; }
    const T& operator[](int i) const { return (&x)[i]; }
};
template <typename T> struct vec4 {
    T x, y, z, w;
    T& operator[](int i) { return (&x)[i]; }
    const T& operator[](int i) const { return (&x)[i]; }
};

template <typename T> inline vec2<T> operator+(vec2<T> a, vec2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> inline vec2<T> operator-(vec2<T> a, vec2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> inline vec2<T> operator*(vec2<T> a, vec2<T> b) { return {a.x * b.x, a.y * b.y}; }
template <typename T> inline vec2<T> operator*(vec2<T> a, T s) { return {a.x * s, a.y * s}; }

template <typename T> inline vec3<T> operator+(vec3<T> a, vec3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline vec3<T> operator-(vec3<T> a, vec3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline vec3<T> operator*(vec3<T> a, vec3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename T> inline vec3<T> operator*(vec3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> inline vec3<T> operator*(T s, vec3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> inline vec3<T> operator+(vec3<T> a, T s) { return {a.x + s, a.y + s, a.z + s}; }
template <typename T> inline vec3<T> operator-(vec3<T> a, T s) { return {a.x - s, a.y - s, a.z - s}; }
template <typename T> inline vec3<T> operator/(vec3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> inline vec3<T> operator-(vec3<T> a) { return {-a.x, -a.y, -a.z}; }

template <typename T> inline vec4<T> operator+(vec4<T> a, vec4<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <typename T> inline vec4<T> operator-(vec4<T> a, vec4<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
template <typename T> inline vec4<T> operator*(vec4<T> a, vec4<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
template <typename T> inline vec4<T> operator*(vec4<T> a, T s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }

// dot: tmp = a*b; (tmp.x + tmp.y) + tmp.z   (func_geometric.inl:48-55)
template <typename T> inline T dot(vec3<T> a, vec3<T> b) {
    const vec3<T> t = a * b;
    return t.x + t.y + t.z;
}
template <typename T> inline vec3<T> normalize(vec3<T> v) { return v * (T(1) / SF<T>::sqrt(dot(v, v))); }
template <typename T> inline vec3<T> vmax(vec3<T> a, vec3<T> b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
template <typename T> inline vec3<T> vmin(vec3<T> a, vec3<T> b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }

// Column-major matrices: m.c[col][row].
template <typename T> struct mat3 {
    vec3<T> c[3];
    vec3<T>& operator[](int i) { return c[i]; }
    const vec3<T>& operator[](int i) const { return c[i]; }
};
template <typename T> struct mat4 {
    vec4<T> c[4];
    vec4<T>& operator[](int i) { return c[i]; }
    const vec4<T>& operator[](int i) const { return c[i]; }
};

template <typename T> inline mat3<T> mat3_from4(const mat4<T>& m) {
    mat3<T> r;
    for (int i = 0; i < 3; ++i) r.c[i] = {m.c[i].x, m.c[i].y, m.c[i].z};
    return r;
}
template <typename T> inline mat3<T> transpose(const mat3<T>& m) {
    mat3<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.c[i][j] = m.c[j][i];
    return r;
}
// m * v = (m0*v.x + m1*v.y) + m2*v.z
template <typename T> inline vec3<T> operator*(const mat3<T>& m, vec3<T> v) {
    return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z;
}
// mul3x3: tmp = A0*B.x; tmp += A1*B.y; tmp += A2*B.z (type_mat3x3.inl:513-542)
template <typename T> inline mat3<T> operator*(const mat3<T>& a, const mat3<T>& b) {
    mat3<T> r;
    for (int j = 0; j < 3; ++j) {
        vec3<T> t = a.c[0] * b.c[j].x;
        t = t + a.c[1] * b.c[j].y;
        t = t + a.c[2] * b.c[j].z;
        r.c[j] = t;
    }
    return r;
}
template <typename T> inline mat3<T> operator-(const mat3<T>& m) { return {{-m.c[0], -m.c[1], -m.c[2]}}; }
// (m0*v0 + m1*v1) + (m2*v2 + m3*v3)   (type_mat4x4.inl:536-560)
template <typename T> inline vec4<T> operator*(const mat4<T>& m, vec4<T> v) {
    const vec4<T> a0 = m.c[0] * v.x + m.c[1] * v.y;
    const vec4<T> a1 = m.c[2] * v.z + m.c[3] * v.w;
    return a0 + a1;
}
// mul4x4 non-SIMD (type_mat4x4.inl:660-700)
template <typename T> inline mat4<T> operator*(const mat4<T>& a, const mat4<T>& b) {
    mat4<T> r;
    for (int j = 0; j < 4; ++j) {
        vec4<T> t = a.c[0] * b.c[j].x;
        t = t + a.c[1] * b.c[j].y;
        t = t + a.c[2] * b.c[j].z;
        t = t + a.c[3] * b.c[j].w;
        r.c[j] = t;
    }
    return r;
}
// inv3x3 non-aligned (func_matrix.inl:353-376)
template <typename T> inline mat3<T> inverse(const mat3<T>& m) {
    const T ood = T(1) / (+m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                          m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                          m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]));
    mat3<T> r;
    r[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]);
    r[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]);
    r[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]);
    r[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]);
    r[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]);
    r[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]);
    r[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]);
    r[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]);
    r[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]);
    for (int i = 0; i < 3; ++i) r.c[i] = r.c[i] * ood;
    return r;
}
// affineInverse (gtc/matrix_inverse.inl:17-26)
template <typename T> inline mat4<T> affine_inverse(const mat4<T>& m) {
    const mat3<T> inv = inverse(mat3_from4(m));
    const vec3<T> t = (-inv) * vec3<T>{m.c[3].x, m.c[3].y, m.c[3].z};
    mat4<T> r;
    for (int i = 0; i < 3; ++i) r.c[i] = {inv.c[i].x, inv.c[i].y, inv.c[i].z, T(0)};
    r.c[3] = {t.x, t.y, t.z, T(1)};
    return r;
}

template <typename T> inline mat4<T> mat4_from(const float* p) {
    mat4<T> m;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) m.c[c][r] = (T)p[4 * c + r];
    return m;
}

}  // namespace orc
