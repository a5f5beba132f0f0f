// gps_fb.h — the 64-bit framebuffer word and its memory layout, shared by the
// gfx950 kernels and host code.
//
// Word  = depth28 << 36 | r12 << 24 | g12 << 12 | b12
//         (src/config.h:36-41, src/core/gaussians/gaussian_types.cuh:533-540)
// depth28 quantizes |view z| linearly between the near and far planes
//         (src/core/math/linalg.cuh:406-411).  The scale factor is the float
//         value of (2^28 - 1), i.e. 2^28, so a depth at the far plane wraps to
//         0 after the 28-bit mask exactly as in the reference.
// Layout: 64x64 Morton-ordered tiles, samples-per-pixel innermost
//         (src/core/rendering/passes/post_processing_kernels.cuh:71-104).
#pragma once

#include "gps_detmath.h"

#define GPS_COLOR_BITS 36ull
#define GPS_DEPTH_BITS 28ull
#define GPS_DEPTH_MAX ((1ull << GPS_DEPTH_BITS) - 1ull)
#define GPS_DEPTH_SHIFT GPS_COLOR_BITS
#define GPS_COLOR_MASK ((1ull << GPS_COLOR_BITS) - 1ull)
#define GPS_DEPTH_MASK (GPS_DEPTH_MAX << GPS_DEPTH_SHIFT)
#define GPS_TILE_BITS 6
#define GPS_TILE_SIZE 64
// float(DEPTH_MAX) as the reference's mixed float * ull arithmetic yields it.
#define GPS_DEPTH_MAX_F 0x1.0p+28f

namespace gps {

GPS_HD uint32_t morton_spread(uint32_t a) {
    a = (a | (a << 8)) & 0x00FF00FFu;
    a = (a | (a << 4)) & 0x0F0F0F0Fu;
    a = (a | (a << 2)) & 0x33333333u;
    a = (a | (a << 1)) & 0x55555555u;
    return a;
}
GPS_HD uint32_t morton_encode(uint32_t x, uint32_t y) { return morton_spread(x) | (morton_spread(y) << 1); }

GPS_HD int32_t fb_tiles_per_row(int32_t width) { return (width + GPS_TILE_SIZE - 1) >> GPS_TILE_BITS; }

// Word index of sample layer `p` of pixel (x, y) in a framebuffer `width`
// pixels wide with `num_passes` layers.
GPS_HD int32_t fb_index(int32_t x, int32_t y, int32_t p, int32_t num_passes, int32_t width) {
    const int32_t tx = x >> GPS_TILE_BITS;
    const int32_t ty = y >> GPS_TILE_BITS;
    const uint32_t code = morton_encode((uint32_t)(x & (GPS_TILE_SIZE - 1)), (uint32_t)(y & (GPS_TILE_SIZE - 1)));
    const int32_t tile = ty * fb_tiles_per_row(width) + tx;
    return (tile * GPS_TILE_SIZE * GPS_TILE_SIZE + (int32_t)code) * num_passes + p;
}

// Number of words of a framebuffer of width x height pixels (padded to whole
// tiles) with num_passes layers.
GPS_HD uint64_t fb_words(int32_t width, int32_t height, int32_t num_passes) {
    const uint64_t tr = (uint64_t)fb_tiles_per_row(width);
    const uint64_t tc = (uint64_t)fb_tiles_per_row(height);
    return tr * tc * (uint64_t)(GPS_TILE_SIZE * GPS_TILE_SIZE) * (uint64_t)num_passes;
}

// view z (negative in front of the camera) -> 28-bit quantized depth.
GPS_HD uint64_t depth_quantize(float view_z, float near_plane, float far_plane) {
    const float t = glm_clamp((det_abs(view_z) - near_plane) / (far_plane - near_plane), 0.0f, 1.0f);
    return det_f2u64_rz(t * GPS_DEPTH_MAX_F) & GPS_DEPTH_MAX;
}
// 28-bit quantized depth -> view z.
GPS_HD float depth_dequantize(uint32_t d28, float near_plane, float far_plane) {
    return -(((float)d28 / GPS_DEPTH_MAX_F) * (far_plane - near_plane) + near_plane);
}
// view z -> NDC depth of a right-handed OpenGL projection (linalg.cuh:413-420).
GPS_HD float view_to_ndc_depth(float z_view, float near_plane, float far_plane) {
    const float A = (far_plane + near_plane) / (far_plane - near_plane);
    const float B = (2.0f * far_plane * near_plane) / (far_plane - near_plane);
    return A + B / z_view;
}

// Colour cache -> packed 3x12-bit colour (gaussian_types.cuh:638-645).
GPS_HD uint64_t pack_color36(float r, float g, float b) {
    const uint32_t r12 = det_f2u_rz(r * 255.0f) & 0xFFFu;
    const uint32_t g12 = det_f2u_rz(g * 255.0f) & 0xFFFu;
    const uint32_t b12 = det_f2u_rz(b * 255.0f) & 0xFFFu;
    return ((uint64_t)r12 << 24) | ((uint64_t)g12 << 12) | (uint64_t)b12;
}

// Background word: all-ones depth | colour (gaussian_point_splatting.cu:620-621,
// sh.cuh:124-133).
GPS_HD uint64_t background_word(float r, float g, float b) {
    const uint64_t c = ((det_f2u64_rz(r * 255.0f) & 0xFFFull) << 24) |
                       ((det_f2u64_rz(g * 255.0f) & 0xFFFull) << 12) |
                       (det_f2u64_rz(b * 255.0f) & 0xFFFull);
    return 0xFFFFFFF000000000ull | c;
}

}  // namespace gps
