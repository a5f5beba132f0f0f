// gps_hostmath.h — host-side camera / matrix math with glm 1.0.2 conventions
// (column-major, right-handed, NDC z in [-1, 1]) and glm's non-SIMD
// evaluation order, so per-frame constants match the reference's
// (gaussian_point_splatting.cu:605-615, render_server.cu:128-129,
// src/world/camera.cpp:88-155).
#pragma once

#include <cmath>
#include <cstring>

namespace gps {
namespace hm {

struct M4 {
    float m[16];  // m[4*col + row]
    float& at(int c, int r) { return m[4 * c + r]; }
    float at(int c, int r) const { return m[4 * c + r]; }
};

inline M4 identity() {
    M4 r;
    std::memset(r.m, 0, sizeof(r.m));
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}

// mul4x4 (glm/detail/type_mat4x4.inl:660-700): col j = ((A0*bx + A1*by) + A2*bz) + A3*bw
inline M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            float t = a.at(0, i) * b.at(j, 0);
            t = t + a.at(1, i) * b.at(j, 1);
            t = t + a.at(2, i) * b.at(j, 2);
            t = t + a.at(3, i) * b.at(j, 3);
            r.at(j, i) = t;
        }
    return r;
}

// affineInverse (glm/gtc/matrix_inverse.inl:17-26) with inv3x3 non-aligned
// (glm/detail/func_matrix.inl:353-376).
inline M4 affine_inverse(const M4& M) {
    float m[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) m[c][r] = M.at(c, r);
    const float ood = 1.0f / (+m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                              m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                              m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]));
    float inv[3][3];
    inv[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]);
    inv[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]);
    inv[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]);
    inv[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]);
    inv[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]);
    inv[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]);
    inv[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]);
    inv[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]);
    inv[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) inv[c][r] = inv[c][r] * ood;
    // -Inv * m[3]: (-c0*x + -c1*y) + -c2*z
    const float x = M.at(3, 0), y = M.at(3, 1), z = M.at(3, 2);
    float t[3];
    for (int r = 0; r < 3; ++r) t[r] = ((-inv[0][r]) * x + (-inv[1][r]) * y) + (-inv[2][r]) * z;
    M4 o;
    for (int c = 0; c < 3; ++c) {
        for (int r = 0; r < 3; ++r) o.at(c, r) = inv[c][r];
        o.at(c, 3) = 0.0f;
    }
    o.at(3, 0) = t[0];
    o.at(3, 1) = t[1];
    o.at(3, 2) = t[2];
    o.at(3, 3) = 1.0f;
    return o;
}

// compute_inverse<4,4> (glm/detail/func_matrix.inl:387-444).
inline M4 inverse(const M4& M) {
    auto m = [&](int c, int r) { return M.at(c, r); };
    const float C00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3);
    const float C02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
    const float C03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
    const float C04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
    const float C06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3);
    const float C07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
    const float C08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2);
    const float C10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
    const float C11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
    const float C12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
    const float C14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3);
    const float C15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
    const float C16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2);
    const float C18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
    const float C19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
    const float C20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
    const float C22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1);
    const float C23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
    const float F0[4] = {C00, C00, C02, C03}, F1[4] = {C04, C04, C06, C07}, F2[4] = {C08, C08, C10, C11};
    const float F3[4] = {C12, C12, C14, C15}, F4[4] = {C16, C16, C18, C19}, F5[4] = {C20, C20, C22, C23};
    const float V0[4] = {m(1, 0), m(0, 0), m(0, 0), m(0, 0)};
    const float V1[4] = {m(1, 1), m(0, 1), m(0, 1), m(0, 1)};
    const float V2[4] = {m(1, 2), m(0, 2), m(0, 2), m(0, 2)};
    const float V3[4] = {m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
    const float SA[4] = {+1, -1, +1, -1}, SB[4] = {-1, +1, -1, +1};
    M4 inv;
    for (int i = 0; i < 4; ++i) {
        inv.at(0, i) = ((V1[i] * F0[i] - V2[i] * F1[i]) + V3[i] * F2[i]) * SA[i];
        inv.at(1, i) = ((V0[i] * F0[i] - V2[i] * F3[i]) + V3[i] * F4[i]) * SB[i];
        inv.at(2, i) = ((V0[i] * F1[i] - V1[i] * F3[i]) + V3[i] * F5[i]) * SA[i];
        inv.at(3, i) = ((V0[i] * F2[i] - V1[i] * F4[i]) + V2[i] * F5[i]) * SB[i];
    }
    const float row0[4] = {inv.at(0, 0), inv.at(1, 0), inv.at(2, 0), inv.at(3, 0)};
    const float d0[4] = {m(0, 0) * row0[0], m(0, 1) * row0[1], m(0, 2) * row0[2], m(0, 3) * row0[3]};
    const float dot1 = (d0[0] + d0[1]) + (d0[2] + d0[3]);
    const float ood = 1.0f / dot1;
    for (int i = 0; i < 16; ++i) inv.m[i] = inv.m[i] * ood;
    return inv;
}

struct V3 {
    float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 x, V3 y) { return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y}; }
inline V3 normalize(V3 v) {
    const float s = 1.0f / std::sqrt(dot(v, v));
    return {v.x * s, v.y * s, v.z * s};
}

// glm::lookAtRH (glm/ext/matrix_transform.inl:153-173)
inline M4 look_at(V3 eye, V3 center, V3 up) {
    const V3 f = normalize(sub(center, eye));
    const V3 s = normalize(cross(f, up));
    const V3 u = cross(s, f);
    M4 r = identity();
    r.at(0, 0) = s.x;
    r.at(1, 0) = s.y;
    r.at(2, 0) = s.z;
    r.at(0, 1) = u.x;
    r.at(1, 1) = u.y;
    r.at(2, 1) = u.z;
    r.at(0, 2) = -f.x;
    r.at(1, 2) = -f.y;
    r.at(2, 2) = -f.z;
    r.at(3, 0) = -dot(s, eye);
    r.at(3, 1) = -dot(u, eye);
    r.at(3, 2) = dot(f, eye);
    return r;
}

// glm::perspectiveRH_NO (glm/ext/matrix_clip_space.inl:249-262) of radians(fovy)
inline M4 perspective(float fovy_deg, float aspect, float zn, float zf) {
    const float fovy = fovy_deg * 0.01745329251994329576923690768489f;
    const float t = std::tan(fovy / 2.0f);
    M4 r;
    std::memset(r.m, 0, sizeof(r.m));
    r.at(0, 0) = 1.0f / (aspect * t);
    r.at(1, 1) = 1.0f / t;
    r.at(2, 2) = -(zf + zn) / (zf - zn);
    r.at(2, 3) = -1.0f;
    r.at(3, 2) = -(2.0f * zf * zn) / (zf - zn);
    return r;
}

// Camera::getFrustumPlanesFromViewMatrix (src/world/camera.cpp:126-155)
inline void frustum_planes(float fovy_deg, const M4& view, float aspect, float zn, float zf, float out[24]) {
    const M4 vp = mul(perspective(fovy_deg, aspect, zn, zf), view);
    // rows of vp = columns of transpose(vp)
    float row[4][4];
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) row[i][k] = vp.at(k, i);
    const int sgn[6] = {+1, -1, +1, -1, +1, -1};
    const int src[6] = {0, 0, 1, 1, 2, 2};
    for (int p = 0; p < 6; ++p) {
        float pl[4];
        for (int k = 0; k < 4; ++k)
            pl[k] = sgn[p] > 0 ? row[3][k] + row[src[p]][k] : row[3][k] - row[src[p]][k];
        const float len = std::sqrt((pl[0] * pl[0] + pl[1] * pl[1]) + pl[2] * pl[2]);
        if (len > 0.0f)
            for (int k = 0; k < 4; ++k) pl[k] = pl[k] / len;
        for (int k = 0; k < 4; ++k) out[4 * p + k] = pl[k];
    }
}

}  // namespace hm
}  // namespace gps
