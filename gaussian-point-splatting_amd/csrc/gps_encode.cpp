// gps_encode.cpp — host encoders of the compact GPS_FORMAT_HALF_FP8 scene
// format: fp16 (IEEE binary16, round-to-nearest-even) rotation / scale /
// opacity / SH DC, and the 45 higher-band SH coefficients as OCP FP8 E4M3
// (bias 7, max 448, no infinities) relative to one fp32 scale per Gaussian,
// scale = max|c| / 448.
//
// SH record (64 bytes): bytes 0-5 DC rgb fp16, 6-7 zero, 8-11 fp32 scale,
// 12-56 the 45 fp8 codes (coefficient-major, rgb interleaved), 57-63 zero.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>
#include <algorithm>

#include "../../include/gps.h"

uint16_t gps_float_to_half_rne(float f);

// float -> E4M3 code, round-to-nearest-even, saturating at +-448.
uint8_t gps_float_to_fp8e4m3_rne(float x) {
    uint32_t bits;
    std::memcpy(&bits, &x, 4);
    const uint8_t sign = (uint8_t)((bits >> 24) & 0x80u);
    const float a = std::fabs(x);
    if (!(a == a)) return (uint8_t)(sign | 0x7Fu);
    if (a == 0.0f) return sign;
    if (a >= 448.0f) return (uint8_t)(sign | 0x7Eu);
    int e;
    std::frexp(a, &e);  // a = f * 2^e, f in [0.5, 1)
    e -= 1;             // a in [2^e, 2^(e+1))
    if (e < -6) {       // subnormal range: units of 2^-9
        const float q = std::nearbyint(a * 512.0f);  // exact scaling, RNE
        return (uint8_t)(sign | (uint8_t)q);         // q == 8 is the smallest normal code
    }
    float q = std::nearbyint(std::ldexp(a, 3 - e));  // in [8, 16]
    if (q >= 16.0f) {
        q = 8.0f;
        e += 1;
    }
    uint32_t code = ((uint32_t)(e + 7) << 3) | ((uint32_t)q - 8u);
    if (code > 0x7Eu) code = 0x7Eu;
    return (uint8_t)(sign | code);
}

void gps_encode_sh_record(const float* sh, int ncoef, uint8_t out[64]) {
    std::memset(out, 0, 64);
    const uint16_t dc[3] = {gps_float_to_half_rne(sh[0]), gps_float_to_half_rne(sh[1]), gps_float_to_half_rne(sh[2])};
    std::memcpy(out, dc, 6);
    const int nrest = 3 * (ncoef - 1);
    float mx = 0.0f;
    for (int k = 0; k < nrest; ++k) mx = std::fmax(mx, std::fabs(sh[3 + k]));
    const float scale = mx > 0.0f ? mx / 448.0f : 1.0f;
    std::memcpy(out + 8, &scale, 4);
    for (int k = 0; k < nrest; ++k) out[12 + k] = gps_float_to_fp8e4m3_rne(sh[3 + k] / scale);
}

extern "C" gps_status gps_encode_half_fp8(const gps_scene_desc* d, uint16_t* rotations, uint16_t* scales,
                                          uint16_t* opacities, uint16_t* sh_dc, uint8_t* sh_rest, float* sh_scale) {
    if (!d || !rotations || !scales || !opacities || !sh_dc || !sh_rest || !sh_scale) return GPS_ERR_INVALID_ARG;
    if (d->sh_degree < 0 || d->sh_degree > 3) return GPS_ERR_INVALID_ARG;
    const int ncoef = (d->sh_degree + 1) * (d->sh_degree + 1);
    // Each Gaussian encodes independently: ranges on host threads.
    auto encode_range = [&](uint64_t g0, uint64_t g1) {
        uint8_t rec[64];
        for (uint64_t g = g0; g < g1; ++g) {
            for (int k = 0; k < 4; ++k) rotations[4 * g + k] = gps_float_to_half_rne(d->rotations[4 * g + k]);
            for (int k = 0; k < 3; ++k) scales[3 * g + k] = gps_float_to_half_rne(d->scales[3 * g + k]);
            opacities[g] = gps_float_to_half_rne(d->opacities[g]);
            gps_encode_sh_record(d->shs + (uint64_t)3 * ncoef * g, ncoef, rec);
            std::memcpy(sh_dc + 3 * g, rec, 6);
            std::memcpy(sh_scale + g, rec + 8, 4);
            std::memset(sh_rest + 45 * g, 0, 45);
            std::memcpy(sh_rest + 45 * g, rec + 12, (size_t)3 * (ncoef - 1));
        }
    };
    const uint64_t n = d->num_gaussians;
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 1 : (nt > 32 ? 32 : nt);
    if (n < 65536) nt = 1;
    const uint64_t chunk = (n + nt - 1) / nt;
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < nt && (uint64_t)t * chunk < n; ++t)
        ts.emplace_back(encode_range, (uint64_t)t * chunk, std::min<uint64_t>(n, (uint64_t)(t + 1) * chunk));
    encode_range(0, std::min<uint64_t>(n, chunk));
    for (auto& t : ts) t.join();
    return GPS_OK;
}

extern "C" uint16_t gps_float_to_half(float f) { return gps_float_to_half_rne(f); }
extern "C" uint8_t gps_float_to_fp8e4m3(float f) { return gps_float_to_fp8e4m3_rne(f); }

// ---------------------------------------------------------------- GPS_FORMAT_QUANTIZED
// The reference's compressed scene (src/core/gaussians/gaussian_types.cuh,
// src/core/math/sh.cuh), encoded on the host in the reference's order of
// operations: scene box :123-170, geometry words :237-320, opacity byte
// :204-207, SH bytes sh.cuh:29-39.  glm's scalar min / max / clamp compare
// with `<` (a NaN survives a clamp); every float -> unsigned conversion here
// is of a value in [0, 2^23], and a NaN (zero-width box) encodes as 0.
namespace {

float q_min(float x, float y) { return (y < x) ? y : x; }
float q_max(float x, float y) { return (x < y) ? y : x; }
float q_clamp(float x, float lo, float hi) { return q_min(q_max(x, lo), hi); }
uint32_t q_trunc(float x) { return x == x ? (uint32_t)x : 0u; }

// quat_fem (gaussian_types.cuh:186-195) of a normalized quaternion with
// q[3] >= 0, in the stored (.x .. .w) order.  2/pi is a double (M_PI), so
// the product is formed in double and rounded once.
void quat_fem(const float q[4], float v[3]) {
    const float w = q[3];
    const float a = 1.0f - w * w;
    const float b = 1.0f / std::sqrt(a + 1e-8f);
    const float k = a * b;
    const float s = (float)((2.0 / 3.14159265358979323846) * (double)std::atan(k / w) * (double)b);
    for (int i = 0; i < 3; ++i) v[i] = s * q[i];
}

}  // namespace

extern "C" gps_status gps_quantized_bounds_compute(const gps_scene_desc* d, gps_quantized_bounds* out) {
    if (!d || !out) return GPS_ERR_INVALID_ARG;
    if (d->sh_degree < 0 || d->sh_degree > 3) return GPS_ERR_INVALID_ARG;
    if (d->num_gaussians > 0 && (!d->means || !d->scales || !d->shs)) return GPS_ERR_INVALID_ARG;
    const float lowest = -3.40282347e38f, highest = 3.40282347e38f;
    for (int c = 0; c < 3; ++c) {
        out->max_pos[c] = lowest;
        out->min_pos[c] = highest;
        out->max_log_scale[c] = lowest;
        out->min_log_scale[c] = highest;
    }
    for (int k = 0; k < 45; ++k) {
        out->max_sh[k] = lowest;
        out->min_sh[k] = highest;
    }
    const int ncoef = (d->sh_degree + 1) * (d->sh_degree + 1);
    for (uint64_t g = 0; g < d->num_gaussians; ++g) {
        for (int c = 0; c < 3; ++c) {
            const float m = d->means[3 * g + c];
            out->max_pos[c] = q_max(out->max_pos[c], m);
            out->min_pos[c] = q_min(out->min_pos[c], m);
            const float ls = std::log(d->scales[3 * g + c]);
            out->max_log_scale[c] = q_max(out->max_log_scale[c], ls);
            out->min_log_scale[c] = q_min(out->min_log_scale[c], ls);
        }
        const float* sh = d->shs + (uint64_t)3 * ncoef * g;
        for (int k = 0; k < 45; ++k) {
            const float v = k < 3 * (ncoef - 1) ? sh[3 + k] : 0.0f;
            out->max_sh[k] = q_max(out->max_sh[k], v);
            out->min_sh[k] = q_min(out->min_sh[k], v);
        }
    }
    for (int c = 0; c < 3; ++c) {
        out->max_pos[c] += 0.001f;
        out->min_pos[c] -= 0.001f;
    }
    return GPS_OK;
}

// One Gaussian: geometry words, opacity byte, SH record (64 bytes: DC fp32
// rgb at 0-11, the 45 codes at 12-56, zero padding).
void gps_encode_quantized_one(const gps_scene_desc* d, const gps_quantized_bounds* b, uint64_t g, uint32_t geo[4],
                              uint8_t* opacity, uint8_t rec[64]) {
    const float* pos = d->means + 3 * g;
    uint32_t p[3];
    const float pmax[3] = {8388607.0f, 4194303.0f, 8388607.0f};  // (1 << 23) - 1, (1 << 22) - 1
    for (int c = 0; c < 3; ++c) {
        const float t = q_clamp((pos[c] - b->min_pos[c]) / (b->max_pos[c] - b->min_pos[c]), 0.0f, 1.0f);
        p[c] = q_trunc(t * pmax[c]);
    }
    // glm::normalize of the stored vec4: v * inversesqrt((x*x + y*y) + (z*z + w*w))
    const float* r = d->rotations + 4 * g;
    const float dot = (r[0] * r[0] + r[1] * r[1]) + (r[2] * r[2] + r[3] * r[3]);
    const float inv = 1.0f / std::sqrt(dot);
    float q[4] = {r[0] * inv, r[1] * inv, r[2] * inv, r[3] * inv};
    if (q[3] < 0.0f)
        for (float& x : q) x = -x;
    float v[3];
    quat_fem(q, v);
    uint32_t rq[3], sq[3];
    for (int c = 0; c < 3; ++c) {
        rq[c] = q_trunc(q_clamp((v[c] + 1.0f) * 0.5f, 0.0f, 1.0f) * 1023.0f);
        const float ls = std::log(d->scales[3 * g + c]);
        const float st = q_clamp((ls - b->min_log_scale[c]) / (b->max_log_scale[c] - b->min_log_scale[c]), 0.0f, 1.0f);
        sq[c] = q_trunc(st * 1023.0f);
    }
    geo[0] = (p[0] & 0x7FFFFFu) | ((p[1] & 0x1FFu) << 23);
    geo[1] = ((p[1] >> 9) & 0x1FFFu) | ((p[2] & 0x7FFFFu) << 13);
    geo[2] = ((p[2] >> 19) & 0xFu) | ((rq[0] & 0x3FFu) << 4) | ((rq[1] & 0x3FFu) << 14) | ((rq[2] & 0xFFu) << 24);
    geo[3] = ((rq[2] >> 8) & 0x3u) | ((sq[0] & 0x3FFu) << 2) | ((sq[1] & 0x3FFu) << 12) | ((sq[2] & 0x3FFu) << 22);
    *opacity = (uint8_t)(std::fmin(std::fmax(d->opacities[g], 0.0f), 1.0f) * 255.0f + 0.5f);
    std::memset(rec, 0, 64);
    const int ncoef = (d->sh_degree + 1) * (d->sh_degree + 1);
    const float* sh = d->shs + (uint64_t)3 * ncoef * g;
    std::memcpy(rec, sh, 12);
    for (int k = 0; k < 45; ++k) {
        const float c = k < 3 * (ncoef - 1) ? sh[3 + k] : 0.0f;
        const float t = (c - b->min_sh[k]) / (b->max_sh[k] - b->min_sh[k]);
        rec[12 + k] = (uint8_t)q_trunc(q_clamp(t * 255.0f, 0.0f, 255.0f));
    }
}

extern "C" gps_status gps_encode_quantized(const gps_scene_desc* d, const gps_quantized_bounds* b, uint32_t* geometry,
                                           uint8_t* opacity, float* sh_dc, uint8_t* sh_codes) {
    if (!d || !b || !geometry || !opacity || !sh_dc || !sh_codes) return GPS_ERR_INVALID_ARG;
    if (d->sh_degree < 0 || d->sh_degree > 3) return GPS_ERR_INVALID_ARG;
    if (d->num_gaussians > 0 && (!d->means || !d->rotations || !d->scales || !d->opacities || !d->shs))
        return GPS_ERR_INVALID_ARG;
    uint8_t rec[64];
    for (uint64_t g = 0; g < d->num_gaussians; ++g) {
        gps_encode_quantized_one(d, b, g, geometry + 4 * g, opacity + g, rec);
        std::memcpy(sh_dc + 3 * g, rec, 12);
        std::memcpy(sh_codes + 45 * g, rec + 12, 45);
    }
    return GPS_OK;
}
