// gps_rng.h — counter-based RNG and the sampling transforms of Gaussian point
// splatting, shared by the gfx950 kernels and host code.
//
// The integer generators are published algorithms used by the reference:
//   hash32      : "lowbias32" integer hash (constants 0x7feb352d, 0x846ca68b),
//                 src/core/random/random.cuh:192-199
//   make_seed   : two hash32 streams mixed with golden-ratio/sqrt constants,
//                 random.cuh:201-215
//   pcg2d       : Jarzynski & Olano, "Hash Functions for GPU Rendering",
//                 JCGT 9(3) 2020, random.cuh:162-184 (u = x * 2^-32)
// The float transforms follow random.cuh:69-141 and poisson.cuh:12-38 with the
// fast-math intrinsics replaced by the deterministic functions of
// gps_detmath.h.  The dilog / inv_dilog polynomial coefficients are the
// reference's own fitted constants (they define the sampling distribution and
// therefore have to be identical).
#pragma once

#include "gps_detmath.h"

namespace gps {

struct u32x2 {
    uint32_t x, y;
};
struct f32x2 {
    float x, y;
};

GPS_HD uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Seed of the stream for work item `idx` in RNG frame `frame`.
GPS_HD u32x2 make_seed(uint32_t idx, uint32_t frame) {
    uint32_t h1 = hash32(idx ^ (frame * 0x9E3779B9u));
    uint32_t h2 = hash32((idx * 0xBB67AE85u) ^ frame);
    h1 ^= h2 * 0x3C6EF372u;
    h2 ^= h1 * 0xA54FF53Au;
    return u32x2{h1, h2};
}

GPS_HD u32x2 pcg2d_uint(u32x2 v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.x += v.y * 1664525u;
    v.y += v.x * 1664525u;
    v.x ^= v.x >> 16;
    v.y ^= v.y >> 16;
    v.x += v.y * 1664525u;
    v.y += v.x * 1664525u;
    v.x ^= v.x >> 16;
    v.y ^= v.y >> 16;
    return v;
}

// Advances the state and returns two uniforms in [0, 1] (1.0 is reachable:
// (float)0xFFFFFFFF rounds to 2^32).
GPS_HD f32x2 pcg2d(u32x2& state) {
    state = pcg2d_uint(state);
    return f32x2{(float)state.x * 0x1.0p-32f, (float)state.y * 0x1.0p-32f};
}

// random.cuh:186-190
GPS_HD int32_t stochastic_round(float x, float u) {
    const int32_t base = det_f2i_rz(det_floor(x));
    const float frac = x - (float)base;
    const int32_t r = (u < frac) ? (int32_t)((uint32_t)base + 1u) : base;
    return r > 0 ? r : 0;
}

// 2*pi as the reference forms it: 2.0f * CUDART_PI_F (3.141592654f).
#define GPS_TWO_PI_F 0x1.921fb6p+2f

// random.cuh:69-75
GPS_HD f32x2 box_muller(float u1, float u2) {
    const float R = det_sqrt(det_fmax(-2.0f * det_logf(glm_clamp(u1, 1.e-37f, 1.0f)), 0.0f));
    const float theta = GPS_TWO_PI_F * u2;
    float s, c;
    det_sincosf(theta, &s, &c);
    return f32x2{R * c, R * s};
}

// Polynomial fit of the dilogarithm Li2 on [0, 1] (random.cuh:88-104).
GPS_HD float dilog(float x) {
    float y = -6.09201442e-01f;
    y = det_fma(y, x, 1.79126616e+00f);
    y = det_fma(y, x, -2.14953223e+00f);
    y = det_fma(y, x, 1.26304372e+00f);
    y = det_fma(y, x, -4.59069895e-01f);
    y = det_fma(y, x, -1.87417414e-01f);
    y = det_fma(y, x, 1.99603130e+00f);
    y = det_fma(y, x, 5.99669467e-05f);
    const float s = 1.0f - x;
    if (s > 0.0f) y += s * det_logf(det_fmax(s, 1e-37f));
    return y;
}

// Li2(1) = pi^2 / 6
#define GPS_LI2_MAX 1.6449340668482264f

// x / GPS_LI2_MAX with IEEE rounding in three operations instead of a full
// division: q0 = x * RN(1/c), the remainder x - q0 * c exactly in one fma,
// one fma correction (Markstein's theorem: with RN(1/c) and q0 within an ulp
// the corrected quotient is the correctly rounded one while the remainder
// does not underflow).  Checked exhaustively against x / c for every
// positive normal float whose quotient is at least 2^-100; smaller, zero
// and NaN inputs take the division.
GPS_HD float det_div_li2max(float x) {
    constexpr float z = 1.0f / GPS_LI2_MAX;
    if (!(det_abs(x) >= 0x1.0p-99f)) return x / GPS_LI2_MAX;
    const float q0 = x * z;
    const float r = det_fma(-q0, GPS_LI2_MAX, x);
    return det_fma(r, z, q0);
}

// The polynomial of inv_dilog at t = min(x / Li2(1), 1).
GPS_HD float inv_dilog_poly(float t) {
    float y = -1.27463503e+01f;
    y = det_fma(y, t, 5.88993459e+01f);
    y = det_fma(y, t, -1.16025780e+02f);
    y = det_fma(y, t, 1.26945827e+02f);
    y = det_fma(y, t, -8.43108826e+01f);
    y = det_fma(y, t, 3.48799862e+01f);
    y = det_fma(y, t, -8.89606235e+00f);
    y = det_fma(y, t, 1.38640936e+00f);
    y = det_fma(y, t, -7.80640876e-01f);
    y = det_fma(y, t, 1.64841888e+00f);
    y = det_fma(y, t, -2.82836687e-05f);
    return y;
}
// Polynomial fit of the inverse of Li2 (random.cuh:110-127).
GPS_HD float inv_dilog(float x) { return inv_dilog_poly(det_fmin(det_div_li2max(x), 1.0f)); }

// Opacity-corrected radial sampling (random.cuh:132-141).  `inv_alpha` is
// 1/max(alpha, 1e-37) and `dilog_alpha` is dilog(alpha); both depend only on
// the Gaussian, so callers may compute them once per Gaussian.
GPS_HD f32x2 corrected_box_muller_pre(float u1, float u2, float inv_alpha, float dilog_alpha) {
    // (1 - u1) * dilog(alpha) is finite and >= 0, so det_fmin's NaN cases
    // cannot arise and a compare-select gives its result; a is finite
    // (inv_alpha <= 1e37), so the clamp is the hardware median of three
    const float q = det_div_li2max((1.0f - u1) * dilog_alpha);
    float a = inv_alpha * inv_dilog_poly(1.0f < q ? 1.0f : q);
#if defined(__HIP_DEVICE_COMPILE__)
    a = __builtin_amdgcn_fmed3f(a, 1.e-37f, 1.0f);
#else
    a = glm_clamp(a, 1.e-37f, 1.0f);
#endif
    // a is normal: det_logf needs no special cases
    // -2 log a is 0 (a = 1) or at least ~1.2e-7 (a <= 1 - 2^-24), at most ~170
    const float R = det_sqrt_normal(det_fmax(-2.0f * det_logf_normal(a), 0.0f));
    const float theta = GPS_TWO_PI_F * u2;  // u2 in [0, 1]: in det_sincosf's range
    float s, c;
    det_sincosf_inrange(theta, &s, &c);
    return f32x2{R * c, R * s};
}
GPS_HD float corrected_inv_alpha(float alpha) { return 1.0f / det_fmax(alpha, 1.e-37f); }

// The alpha test of a sample (gaussian_point_splatting.cu:118-123):
// gw = sigma > 0 ? 0 : exp(sigma), rejected when opacity * gw < 1/255.
GPS_HD bool alpha_test(float opacity, float sigma) {
    const float gw = sigma > 0.0f ? 0.0f : det_expf(sigma);
    return !(opacity * gw < 1.0f / 255.0f);
}
// The exponent at which the test flips: t = ln((1/255) / opacity).  The
// test is monotone in sigma; its decision is fixed outside
// [t - GPS_ALPHA_BAND, t + GPS_ALPHA_BAND], a band far wider than the
// combined error of det_expf (a few ulps), the product's rounding and t's
// own (det_logf of a rounded quotient): about 2e-5 in the exponent for
// |t| < 100.  Subnormal and zero exponentials lie far below the band.
#define GPS_ALPHA_BAND 0x1.0p-8f
GPS_HD float alpha_threshold(float opacity) { return det_logf((1.0f / 255.0f) / opacity); }
// alpha_test's decision, with the exponential evaluated only inside the band
// (t = alpha_threshold(opacity); NaN anywhere falls through to alpha_test).
GPS_HD bool alpha_test_banded(float opacity, float t, float sigma, float band = GPS_ALPHA_BAND) {
    if (sigma < t - band) return false;
    if (sigma > t + band && !(sigma > 0.0f)) return true;
    return alpha_test(opacity, sigma);
}
GPS_HD f32x2 corrected_box_muller(float u1, float u2, float alpha) {
    return corrected_box_muller_pre(u1, u2, corrected_inv_alpha(alpha), dilog(alpha));
}

// Giles' QN3 normal-asymptotic Poisson sampler, M.B. Giles, "Algorithm 955",
// ACM TOMS 42(1), 2016 (poisson.cuh:12-38).  Consumes one pcg2d draw.
GPS_HD uint32_t poisson(u32x2& state, float lambda) {
    const f32x2 u = pcg2d(state);
    const float w = box_muller(u.x, u.y).x;
    const float w2 = w * w;
    const float w3 = w2 * w;
    const float w4 = w2 * w2;
    const float s = det_sqrt(lambda);
    const float inv_s = 1.0f / s;
    const float inv_l = 1.0f / lambda;
    float kf = lambda + s * w + (w2 - 1.0f) / 6.0f;
    kf += inv_s * (-(1.0f / 36.0f) * w - (1.0f / 72.0f) * w3);
    kf += inv_l * (-(8.0f / 405.0f) + (7.0f / 810.0f) * w2 + (1.0f / 270.0f) * w4);
    const int32_t ki = det_f2i_rn(kf);
    return (uint32_t)(ki > 0 ? ki : 0);
}

}  // namespace gps
