// gps_detmath.h — deterministic fp32 scalar math shared by the gfx950 kernels
// (hipcc) and the host code (g++ / clang).
//
// Every function here is built only from IEEE-754 basic operations that are
// correctly rounded on both targets (+, -, *, /, sqrt, fma, floor, ceil, rint)
// plus integer/bit manipulation.  Compiled with -ffp-contract=off and without
// fast-math, the same input therefore yields the same output bits on an x86-64
// host and on an MI355X, which is what makes GPU <-> CPU bit-exact parity of
// sample positions possible.
//
// The reference renderer calls CUDA fast-math intrinsics for the same purposes
// (__logf, __expf, __sincosf, __sinf/__cosf, __float2int_rn, rsqrtf; see
// src/core/random/random.cuh:70,101,135,139, src/core/math/linalg.cuh:316,
// src/core/random/poisson.cuh:34, src/core/gaussians/gaussian_types.cuh:222-225).
// Those are not reproducible off NVIDIA hardware; these are project-written
// replacements from named textbook methods:
//   exp    : Cody-Waite reduction x = k*ln2 + r, |r| <= ln2/2, degree-7 Taylor
//            polynomial (Horner, fma), exact 2^k scaling in two steps.
//   log    : exponent/mantissa split with m in [sqrt(1/2), sqrt(2)],
//            log(m) = 2*atanh(s), s = (m-1)/(m+1), odd series to s^11.
//   sincos : Cody-Waite reduction by pi/2 (three-part constant, fma),
//            Taylor polynomials on [-pi/4, pi/4] (sin to r^9, cos to r^10),
//            quadrant selection.
// Accuracy (checked against FP64 libm in tests/test_detmath.py): <= 2 ulp
// over the ranges the renderer uses.
//
// Float->int conversions follow the documented saturating semantics of the
// PTX cvt instructions the reference compiles to (NaN -> 0, out-of-range ->
// INT_MIN/INT_MAX or 0/UINT_MAX), because plain C++ casts are undefined there
// and x86 / AMDGPU disagree.
#pragma once

#include <stdint.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define GPS_HD __host__ __device__ __forceinline__
#define GPS_ON_DEVICE_COMPILER 1
#else
#define GPS_HD static inline
#endif

#if !defined(__HIP_DEVICE_COMPILE__)
#include <math.h>
#endif

namespace gps {

GPS_HD uint32_t f2u_bits(float x) { return __builtin_bit_cast(uint32_t, x); }
GPS_HD float u2f_bits(uint32_t u) { return __builtin_bit_cast(float, u); }

GPS_HD float det_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
GPS_HD float det_sqrt(float x) { return __builtin_sqrtf(x); }
// det_sqrt for x = 0, x >= 2^-96 (up to infinity), negative x or NaN (both
// NaN): the hardware square root and the +-1-ulp fma correction of the
// IEEE-rounded sequence, without its scaling of tiny inputs and its
// special-case selects, which change nothing there.  The GPU test
// test_device_sqrt_range_exhaustive compares it with det_sqrt for every
// float in [1, 4) (every mantissa at both exponent parities; the rounded
// root commutes with exact scaling by 4), a sample of 2^-96 .. FLT_MAX and
// the special values.  The host takes det_sqrt.
GPS_HD float det_sqrt_normal(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float y = __builtin_amdgcn_sqrtf(x);
    const float ym = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, y) - 1u);
    const float yp = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, y) + 1u);
    const float rm = __builtin_fmaf(-ym, y, x);
    const float rp = __builtin_fmaf(-yp, y, x);
    float r = rm <= 0.0f ? ym : y;
    r = rp > 0.0f ? yp : r;
    return x == 0.0f ? 0.0f : r;
#else
    return det_sqrt(x);
#endif
}
GPS_HD float det_floor(float x) { return __builtin_floorf(x); }
GPS_HD float det_ceil(float x) { return __builtin_ceilf(x); }
GPS_HD float det_rint(float x) { return __builtin_rintf(x); }
GPS_HD float det_abs(float x) { return __builtin_fabsf(x); }

// C99 fminf/fmaxf semantics (a NaN operand yields the other operand) with a
// fixed choice on ties (first operand), so signed zeros resolve identically on
// both targets.
GPS_HD float det_fmin(float a, float b) {
    if (b != b) return a;
    if (a != a) return b;
    return (b < a) ? b : a;
}
GPS_HD float det_fmax(float a, float b) {
    if (b != b) return a;
    if (a != a) return b;
    return (a < b) ? b : a;
}
// glm::min / glm::max / glm::clamp on scalars (glm/detail/func_common.inl:17-35,649-653).
GPS_HD float glm_min(float x, float y) { return (y < x) ? y : x; }
GPS_HD float glm_max(float x, float y) { return (x < y) ? y : x; }
GPS_HD float glm_clamp(float x, float lo, float hi) { return glm_min(glm_max(x, lo), hi); }

// Saturating conversions (PTX cvt.rzi / cvt.rni semantics).  det_f2i_rz is
// written with selects (no branches in device code): the conversion is of a
// value clamped into range, and the saturated / NaN results are selected.
GPS_HD int32_t det_f2i_rz(float x) {
    const float c = (x != x) ? 0.0f : glm_min(glm_max(x, -2147483648.0f), 2147483520.0f);
    const int32_t r = (int32_t)c;  // 2147483520 = the largest float below 2^31
    return (x >= 2147483648.0f) ? 2147483647 : r;
}
GPS_HD int32_t det_f2i_rn(float x) { return det_f2i_rz(det_rint(x)); }
// det_f2i_rz(v) where only its order against lo and hi matters (the caller
// clamps it into [lo, hi] or rejects beyond them; lo <= 0 <= hi, small
// integers): NaN reads as 0 as there, the float is clamped, then converted.
GPS_HD int32_t det_f2i_clamped(float v, float lo, float hi) {
    const float s = (v != v) ? 0.0f : glm_min(glm_max(v, lo), hi);
    return (int32_t)s;
}
// (float)det_f2i_rz(v) for integral-valued v >= 0 or NaN: v below 2^31, the
// saturated 2^31 from there, 0 for NaN.
GPS_HD float det_f2i_rz_float(float v) {
    return (v != v) ? 0.0f : glm_min(v, 2147483648.0f);
}

// Whether det_f2i_rz(v) lies in [lo, hi) for integral-valued (floored) v and
// 0 <= lo < hi <= 65536, and that value in *p: NaN converts to 0, and
// infinities and values beyond the int range convert to values outside such
// bounds, so the test is two float compares of v with NaN read as 0.
GPS_HD bool det_pixel_in(float v, int32_t lo, int32_t hi, int32_t* p) {
    const float s = (v != v) ? 0.0f : v;
    const bool in = s >= (float)lo && s < (float)hi;
    *p = in ? (int32_t)s : 0;
    return in;
}
GPS_HD uint32_t det_f2u_rz(float x) {
    if (!(x > 0.0f)) return 0u;  // NaN, negative, zero
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
GPS_HD uint64_t det_f2u64_rz(float x) {
    if (!(x > 0.0f)) return 0ull;
    if (x >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)x;
}

// ---------------------------------------------------------------- exp
// The special cases (NaN, overflow above ln(FLT_MAX), underflow below
// ln(min subnormal / 2)) are selected after evaluating an in-range stand-in,
// so device code has no branches; in-range inputs take the same arithmetic.
GPS_HD float det_expf(float x) {
    const bool nan = x != x;
    const bool over = x > 0x1.62e430p+6f;
    const bool under = x < -0x1.9fe368p+6f;
    const float xs = (nan || over || under) ? 0.0f : x;
    const float k = det_rint(xs * 0x1.715476p+0f);          // x / ln2
    float r = det_fma(-k, 0x1.62e400p-1f, xs);               // ln2 hi
    r = det_fma(-k, 0x1.7f7d1cp-20f, r);                     // ln2 lo
    float p = 0x1.a01a02p-13f;                               // 1/7!
    p = det_fma(p, r, 0x1.6c16c2p-10f);                      // 1/6!
    p = det_fma(p, r, 0x1.111112p-7f);                       // 1/5!
    p = det_fma(p, r, 0x1.555556p-5f);                       // 1/4!
    p = det_fma(p, r, 0x1.555556p-3f);                       // 1/3!
    p = det_fma(p, r, 0x1.000000p-1f);                       // 1/2!
    p = det_fma(p, r, 1.0f);
    p = det_fma(p, r, 1.0f);
    const int32_t ki = (int32_t)k;                           // |k| <= 150
    const int32_t k1 = ki / 2;
    const int32_t k2 = ki - k1;
    const float s1 = u2f_bits((uint32_t)(k1 + 127) << 23);
    const float s2 = u2f_bits((uint32_t)(k2 + 127) << 23);
    const float e = (p * s1) * s2;
    return nan ? x : over ? u2f_bits(0x7F800000u) : under ? 0.0f : e;
}

// ---------------------------------------------------------------- log
// f / (2 + f), IEEE-rounded, for the f of det_logf_normal (m - 1 with m a
// float in (sqrt(2)/2, sqrt(2)]: 2^23 values).  The device takes the
// hardware reciprocal and one fma correction of the quotient (q0 = f * y,
// the remainder f - q0 * d exact in one fma, q0 + r * y); the GPU test
// test_device_log_quotient_exhaustive compares it with the division for
// every one of those 2^23 inputs.  The host divides.
GPS_HD float det_log_quotient(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float d = 2.0f + f;
    const float y = __builtin_amdgcn_rcpf(d);
    const float q0 = f * y;
    const float r = __builtin_fmaf(-q0, d, f);
    return __builtin_fmaf(r, y, q0);
#else
    return f / (2.0f + f);
#endif
}

// det_logf for a positive normal finite x (the special cases of det_logf
// cannot occur); same arithmetic, so the same bits for such x.
GPS_HD float det_logf_normal(float x, int32_t e = 0) {
    const uint32_t ix = f2u_bits(x);
    e += (int32_t)(ix >> 23) - 127;
    float m = u2f_bits((ix & 0x007FFFFFu) | 0x3F800000u);    // [1, 2)
    if (m > 0x1.6a09e6p+0f) {                                // > sqrt(2)
        m = m * 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;                                // exact
    const float s = det_log_quotient(f);
    const float z = s * s;
    float p = 0x1.745d18p-4f;                                // 1/11
    p = det_fma(p, z, 0x1.c71c72p-4f);                       // 1/9
    p = det_fma(p, z, 0x1.24924ap-3f);                       // 1/7
    p = det_fma(p, z, 0x1.99999ap-3f);                       // 1/5
    p = det_fma(p, z, 0x1.555556p-2f);                       // 1/3
    p = p * z;
    const float t = s + s;                                   // exact
    const float logm = det_fma(t, p, t);
    const float fe = (float)e;
    return det_fma(fe, 0x1.62e400p-1f, det_fma(fe, 0x1.7f7d1cp-20f, logm));
}

GPS_HD float det_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return u2f_bits(0x7FC00000u);
    if (x == 0.0f) return u2f_bits(0xFF800000u);
    if (x == u2f_bits(0x7F800000u)) return x;
    if (x < 0x1.0p-126f) return det_logf_normal(x * 0x1.0p+25f, -25);  // subnormal: exact rescale
    return det_logf_normal(x);
}

// ---------------------------------------------------------------- sin/cos
// For finite |x| < 2^20 (det_sincosf's valid range); same bits as det_sincosf.
GPS_HD void det_sincosf_inrange(float x, float* sp, float* cp) {
    const float k = det_rint(x * 0x1.45f306p-1f);            // x * 2/pi
    float r = det_fma(-k, 0x1.921fb6p+0f, x);
    r = det_fma(-k, -0x1.777a5cp-25f, r);
    r = det_fma(-k, -0x1.000000p-49f, r);
    const float z = r * r;
    float ps = 0x1.71de3ap-19f;                              //  1/9!
    ps = det_fma(ps, z, -0x1.a01a02p-13f);                   // -1/7!
    ps = det_fma(ps, z, 0x1.111112p-7f);                     //  1/5!
    ps = det_fma(ps, z, -0x1.555556p-3f);                    // -1/3!
    const float sr = det_fma(r * z, ps, r);
    float pc = -0x1.27e4fcp-22f;                             // -1/10!
    pc = det_fma(pc, z, 0x1.a01a02p-16f);                    //  1/8!
    pc = det_fma(pc, z, -0x1.6c16c2p-10f);                   // -1/6!
    pc = det_fma(pc, z, 0x1.555556p-5f);                     //  1/4!
    pc = det_fma(pc, z, -0x1.000000p-1f);                    // -1/2!
    const float cr = det_fma(pc, z, 1.0f);
    // quadrant q: (sin, cos) = (sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr),
    // as selects (no branches in device code)
    const uint32_t q = (uint32_t)(int32_t)k & 3u;
    const bool odd = (q & 1u) != 0u;
    const float s = odd ? cr : sr;
    const float c = odd ? sr : cr;
    *sp = (q & 2u) ? -s : s;
    *cp = ((q + 1u) & 2u) ? -c : c;
}

GPS_HD void det_sincosf(float x, float* sp, float* cp) {
    if (!(det_abs(x) < 0x1.0p+20f)) {  // NaN/inf/huge: outside the renderer's use
        const float n = u2f_bits(0x7FC00000u);
        *sp = n;
        *cp = n;
        return;
    }
    det_sincosf_inrange(x, sp, cp);
}

GPS_HD float det_sinf(float x) { float s, c; det_sincosf(x, &s, &c); return s; }
GPS_HD float det_cosf(float x) { float s, c; det_sincosf(x, &s, &c); return c; }

// glm::inversesqrt(x) = 1 / sqrt(x) (glm/detail/func_exponential.inl).
GPS_HD float det_inversesqrt(float x) { return 1.0f / det_sqrt(x); }

}  // namespace gps
