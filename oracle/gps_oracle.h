/*
 * gps_oracle.h — C API of the CPU oracle (plain C++ transcription of the
 * reference's per-frame hot path).
 *
 * TEST INFRASTRUCTURE.  Only tests/, bench.py's CPU-baseline leg and the
 * smoke check load this library, and only as the checker / CPU baseline; the
 * product never links it.
 *
 * The oracle implements the same POD structs as include/gps.h so a test can
 * drive both with identical inputs.
 */
#ifndef GPS_ORACLE_H
#define GPS_ORACLE_H

#include <stdint.h>

#include "../include/gps.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_pass orc_pass;

/* Encoded GPS_FORMAT_HALF_FP8 scene as produced by gps_encode_half_fp8(). */
typedef struct orc_half_fp8_scene {
    uint64_t num_gaussians;
    int32_t sh_degree;
    int32_t reserved;
    const float* means;        /* N*3 */
    const uint16_t* rotations; /* N*4 fp16 (w,x,y,z) */
    const uint16_t* scales;    /* N*3 fp16 */
    const uint16_t* opacities; /* N   fp16 */
    const uint16_t* sh_dc;     /* N*3 fp16 */
    const uint8_t* sh_rest;    /* N*45 fp8 E4M3 */
    const float* sh_scale;     /* N */
} orc_half_fp8_scene;

/* Encoded GPS_FORMAT_QUANTIZED scene as produced by gps_encode_quantized(). */
typedef struct orc_quantized_scene {
    uint64_t num_gaussians;
    gps_quantized_bounds bounds;
    const uint32_t* geometry; /* N*4 */
    const uint8_t* opacity;   /* N */
    const float* sh_dc;       /* N*3 */
    const uint8_t* sh_codes;  /* N*45 */
} orc_quantized_scene;

orc_pass* orc_pass_create(const gps_render_settings* settings, const gps_scene_desc* scene, int32_t nthreads);
orc_pass* orc_pass_create_quantized(const gps_render_settings* settings, const orc_quantized_scene* scene,
                                    int32_t nthreads);
/* Decoded Gaussian g of a quantized scene: mean[3], rotation[4] (stored
 * order), scale[3], opacity, 16 SH coefficients x rgb (59 floats). */
void orc_quantized_decode(const orc_quantized_scene* scene, uint64_t g, float* out);
orc_pass* orc_pass_create_half_fp8(const gps_render_settings* settings, const orc_half_fp8_scene* scene,
                                   int32_t nthreads);
void orc_pass_destroy(orc_pass* pass);
void orc_pass_set_threads(orc_pass* pass, int32_t nthreads);
int32_t orc_pass_render(orc_pass* pass, const gps_frame_params* frame, float* color, float* depth);
int32_t orc_pass_stats(orc_pass* pass, gps_stats* out);
uint64_t orc_pass_debug_size(orc_pass* pass, int32_t which);
int32_t orc_pass_debug_copy(orc_pass* pass, int32_t which, void* dst, uint64_t count);
/* Biased-splat events of the last render (use_unbiased_2d_splatting = 0):
 * out[0] = samples of packed colour 0 that passed the pre-check, out[1] =
 * equal-depth adds that carried out of the 36 colour bits.  When both are 0
 * every framebuffer word is independent of the order of its samples. */
void orc_pass_biased_events(orc_pass* pass, uint64_t out[2]);
/* Per render pass of the last frame (second pass zero without occlusion
 * culling), why each Gaussian of an unculled block did not become valid, in
 * the order of GPS_DBG_REJECTS: tested, valid in the first pass, opacity,
 * projection, off screen, small, Hi-Z, count 0, valid. */
void orc_pass_rejects(orc_pass* pass, uint64_t out[18]);
/* One biased sample on one word: precheck != 0 applies the splat's pre-check
 * and atomicAddColor(add_if_equal = true) (gaussian_point_splatting.cu:130-138);
 * precheck == 0 applies atomicAddColor alone with the given add_if_equal and
 * returns its result (post_processing_kernels.cuh:13-42). */
uint64_t orc_biased_sample(uint64_t* word, uint64_t value, int32_t precheck, int32_t add_if_equal);
/* Per-stage wall times (ms) of the last render: clear, cull, preprocess,
 * distribute, splat, combine, hiz. */
void orc_pass_stage_ms(orc_pass* pass, double out[8]);

/* Per-Gaussian projection in FP32 or FP64 for the accuracy checks.  Writes
 * N x 16 doubles: ok, view_z, pixel_mean.x, pixel_mean.y, cov2d.a, cov2d.b,
 * cov2d.c, color.r, color.g, color.b, importance, ndc_z, bounds min x/y,
 * max x/y (FP32 only).  ok = 1 if the
 * Gaussian passes projection + bounds; colour is evaluated for every such
 * Gaussian. */
/* Carried frame state, for rendering frame f without rendering frames
 * 0 .. f-1 first: the Hi-Z chain (GPS_DBG_HIZ layout, coarsest level first)
 * and the valid mask that the next frame's first pass does not overwrite
 * (its second pass reads it as the previous mask when the first pass yields
 * no points; workload_distributor.cu:32-40,111-115).  Nothing else survives
 * a frame: the framebuffer is cleared, the weights and the first pass's mask
 * are rewritten. */
int32_t orc_pass_set_state(orc_pass* pass, const float* hiz, uint64_t hiz_count, const uint32_t* mask_other,
                           uint64_t mask_words);
/* Block culling alone (compute_occlusion_mask) against the Hi-Z chain the
 * pass currently holds; the mask is then read with GPS_DBG_CULL_MASK. */
int32_t orc_pass_cull(orc_pass* pass, const gps_frame_params* frame);
int32_t orc_project(orc_pass* pass, const gps_frame_params* frame, int32_t use_fp64, double* out);
/* orc_project's FP32 values for the Gaussians `ids` only (n x 16 doubles). */
int32_t orc_project_ids(orc_pass* pass, const gps_frame_params* frame, const uint32_t* ids, uint64_t n, double* out);

/* Resolve kernels (post_processing_kernels.cuh). */
void orc_resolve_taa(const float* cur_color, const float* cur_depth, const float* prev_color,
                     const float* prev_depth, const float reprojection[16], float blend_factor, int32_t width,
                     int32_t height, int32_t frame_number, float* out_color, uint8_t* out_rgb8);
void orc_resolve_progressive(float* color, float* accum, int32_t width, int32_t height, int32_t num_frames,
                             uint8_t* out_rgb8);

/* Scalar probes for known-answer tests. */
void orc_rng_probe(uint32_t idx, uint32_t frame, int32_t draws, uint32_t* out_states /* 2 + 2*draws */);
float orc_dilog(float x);
float orc_inv_dilog(float x);
void orc_corrected_box_muller(float u1, float u2, float alpha, float* out2);
uint32_t orc_poisson(uint32_t s0, uint32_t s1, float lambda);
int32_t orc_stochastic_round(float x, float u);
int32_t orc_pos_to_memory(int32_t x, int32_t y, int32_t p, int32_t num_passes, int32_t width, int32_t height);
uint64_t orc_depth_bits(float view_z, float near_plane, float far_plane);
/* Shared deterministic math as the oracle evaluates it (kinds as gps_selftest_math). */
void orc_detmath(int32_t kind, const float* in, float* out, uint64_t n);
/* computeColorFromSHFloat (full degree 3) at direction dir, FP32 or FP64. */
void orc_sh_eval(const float dir[3], const float* coeffs48, int32_t fp64, double out[3]);
void orc_sh_eval_compact(const float dir[3], const float* coeffs48, int32_t fp64, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
