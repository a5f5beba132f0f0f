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
    uint8_t rec[64];
    for (uint64_t g = 0; g < d->num_gaussians; ++g) {
        for (int k = 0; k < 4; ++k) rotations[4 * g + k] = gps_float_to_half_rne(d->rotations[4 * g + k]);
        for (int k = 0; k < 3; ++k) scales[3 * g + k] = gps_float_to_half_rne(d->scales[3 * g + k]);
        opacities[g] = gps_float_to_half_rne(d->opacities[g]);
        gps_encode_sh_record(d->shs + (uint64_t)3 * ncoef * g, ncoef, rec);
        std::memcpy(sh_dc + 3 * g, rec, 6);
        std::memcpy(sh_scale + g, rec + 8, 4);
        std::memset(sh_rest + 45 * g, 0, 45);
        std::memcpy(sh_rest + 45 * g, rec + 12, (size_t)3 * (ncoef - 1));
    }
    return GPS_OK;
}

extern "C" uint16_t gps_float_to_half(float f) { return gps_float_to_half_rne(f); }
extern "C" uint8_t gps_float_to_fp8e4m3(float f) { return gps_float_to_fp8e4m3_rne(f); }
