/*
 * gps.h — C ABI of the MI355X-native Gaussian point splatting hot path.
 *
 * One frame of the reference renderer is
 *   clear -> [block cull] -> preprocess -> workload distribution -> splat
 *   -> [depth combine -> Hi-Z -> second pass] -> combine -> [Hi-Z]
 *   -> TAA / progressive resolve
 * (reference: src/core/rendering/passes/gaussian_point_splatting.cu:599-737 and
 *  src/core/rendering/render_server.cu:86-157).  This header exposes that path
 * as plain C: POD structs, opaque handles, device pointers as void*, status
 * codes instead of exceptions.  Streams are hipStream_t passed as void*
 * (NULL = the null stream).  No call inside gps_pass_render() or gps_resolve()
 * synchronizes the host, so both can be captured into a hipGraph.
 *
 * Every entry point names the reference interface it replaces.
 */
#ifndef GPS_H
#define GPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPS_ABI_VERSION 3

typedef enum gps_status {
    GPS_OK = 0,
    GPS_ERR_INVALID_ARG = 1,
    GPS_ERR_OUT_OF_MEMORY = 2,
    GPS_ERR_HIP = 3,
    GPS_ERR_UNSUPPORTED = 4,
    GPS_ERR_IO = 5,
    GPS_ERR_NOT_READY = 6
} gps_status;

/* Storage format of the resident scene. */
typedef enum gps_scene_format {
    /* fp32 everything; the reference's GPUGaussianSceneUncompressed
       (src/core/gaussians/gaussian_types.cuh:548-646). */
    GPS_FORMAT_FP32 = 0,
    /* fp32 positions, fp16 scale / rotation / opacity, fp8-E4M3 SH with one
       fp32 scale per Gaussian.  Project-defined compact format. */
    GPS_FORMAT_HALF_FP8 = 1,
    /* The reference's compressed scene, its default build
       (USE_COMPRESSED_GAUSSIANS, src/config.h:15; GPUGaussianSceneCompressed,
       src/core/gaussians/gaussian_types.cuh:421-541): 16-byte geometry
       (position 23/22/23 bits in the scene box, rotation in three 10-bit
       "fem" components, log-scale 10 bits each in the scene's log-scale box;
       encode / decode :237-374, quat_iem :212-235), one opacity byte
       (/255, :199-202), SH DC in fp32 and the 45 higher coefficients as
       bytes in per-coefficient scene ranges (CompactSH,
       src/core/math/sh.cuh:28-51).  Colour is evaluated with the Sloan
       degree-3 basis whatever sh_degree is (computeColorFromCompactSH,
       sh.cuh:257-273 with eval_sh_sloan_l3 :55-96). */
    GPS_FORMAT_QUANTIZED = 2
} gps_scene_format;

/* Host-side scene description, one record per Gaussian in structure-of-arrays
 * form.  Mirrors CPUGaussian (src/core/gaussians/gaussian_types.cuh:89-100):
 *   means      N*3  world position
 *   rotations  N*4  quaternion stored (w, x, y, z)  (gaussian_loader.cu:214-220)
 *   scales     N*3  linear (post-exp) scale
 *   opacities  N    post-sigmoid opacity
 *   shs        N*3*(sh_degree+1)^2  3DGS SH layout: coefficient-major, RGB
 *              interleaved (sh[3*i + c]); coefficients above sh_degree are 0.
 * The arrays are copied; the caller keeps ownership. */
typedef struct gps_scene_desc {
    uint64_t num_gaussians;
    int32_t sh_degree; /* 0..3 */
    int32_t format;    /* gps_scene_format */
    const float* means;
    const float* rotations;
    const float* scales;
    const float* opacities;
    const float* shs;
} gps_scene_desc;

typedef struct gps_scene gps_scene;
typedef struct gps_pass gps_pass;

/* Create-time settings of a pass; the fields of RenderSettings the
 * GaussianPointSplatting constructor reads (gaussian_point_splatting.cu:777-816,
 * src/render_settings.cuh:36-88). */
typedef struct gps_render_settings {
    int32_t width, height;               /* output resolution (Hi-Z size) */
    int32_t render_pass_count;           /* samples per pixel ("spp"), >= 1 */
    int32_t supersampling_factor;        /* >= 1 */
    int64_t max_number_of_points;        /* point budget (default 250e6) */
    int32_t fully_disable_hierarchical_culling;
    int32_t gaussians_per_block;         /* culling block size; 0 = reference
                                            rule max(4, floor(N / 1e6)) */
    int32_t flexible_frame_size;         /* 0: every frame is width x height with
                                            render_pass_count spp, as in the
                                            reference (main.cu:82-86); 1: frames
                                            up to that size and spp are accepted
                                            (serving); occlusion culling still
                                            needs the full size */
} gps_render_settings;

/* Per-frame inputs; the RenderSettings fields read by
 * GaussianPointSplatting::render (gaussian_point_splatting.cu:599-737).
 * Matrices are column-major 4x4 (glm layout). */
typedef struct gps_frame_params {
    float view[16];            /* cameraViewMatrix */
    float proj[16];            /* cameraProjectionMatrix */
    float frustum_planes[24];  /* 6 normalized planes (a,b,c,d): L,R,B,T,N,F */
    float fovy_degrees;
    float near_plane, far_plane;
    float background[3];
    int32_t width, height;
    int32_t supersampling_factor;
    int32_t frame;             /* settings.frame; RNG frames are 4*frame+{0..3} */
    int32_t enable_occlusion_culling;
    int32_t enable_hierarchical_culling;
    int32_t cull_small_gaussians;
    int32_t use_unbiased_2d_splatting;
    int32_t reduce_point_count;
    int32_t occlusion_culling_min_reduce_count;
    int32_t samples_per_pixel; /* 0 = the pass's render_pass_count; other values
                                  need flexible_frame_size */
} gps_frame_params;

/* Counters of the last rendered frame (Statistics, render_settings.cuh:23-34,
 * plus the work units the benchmark reports). */
typedef struct gps_stats {
    uint64_t num_gaussians;
    uint64_t valid_gaussians[2];   /* per render pass (pass 2 only with occlusion culling) */
    uint64_t total_weight[2];      /* uncapped sum of point counts per pass */
    uint64_t num_points[2];        /* points splatted per pass (after the budget cap) */
    uint64_t samples;              /* sum over passes of points * K * spp */
    uint64_t culled_blocks[2];
    int32_t passes;                /* 1 or 2 */
    int32_t budget_truncated;      /* 1 if the point budget cut points */
    int32_t K;                     /* samples per point per layer */
    int32_t gaussians_per_block;   /* culling block size in use (0 when block culling is off) */
    uint64_t atomics;              /* framebuffer atomicMin operations issued (diagnostic) */
    uint64_t num_blocks;           /* culling blocks (0 when block culling is off) */
    int32_t splat_binned;          /* 1: the last pass used the binned splat */
    int32_t splat_fallback;        /* the point-parallel splat ran instead: 1 region lists overflowed,
                                      2 more than 8 samples per framebuffer word (GPS_SPLAT_AUTO) */
    uint32_t bin_entries;          /* (Gaussian, region) pairs of the last pass */
    uint32_t bin_jobs;             /* region jobs of the last pass */
    uint32_t bin_big;              /* Gaussians splatted with global atomics (too many regions) */
    uint32_t reserved;
    uint64_t culled_groups[2];     /* per pass: 64-Gaussian groups the group occlusion test skipped
                                      (counted with gps_pass_set_counters flags on, else 0) */
} gps_stats;

/* -------------------------------------------------------------- versioning */
int32_t gps_abi_version(void);
const char* gps_status_str(gps_status s);
/* Last HIP error string recorded by the library in this thread. */
const char* gps_last_error(void);

/* ------------------------------------------------------------------ scenes */
/* Uploads a scene into HBM.  Replaces GPUGaussianScene*::upload_from_cpu
 * (gaussian_types.cuh:442-503, 568-608). */
gps_status gps_scene_create(const gps_scene_desc* desc, gps_scene** out);
/* Chunked upload of large scenes (the reference uploads in 1M chunks,
 * gaussian_types.cuh:454-501): create an empty scene of `num_gaussians`
 * (desc arrays unused), then fill index ranges with gps_scene_upload_range,
 * whose desc holds `count` Gaussians for [first, first + count). */
gps_status gps_scene_create_empty(uint64_t num_gaussians, int32_t sh_degree, int32_t format, gps_scene** out);
gps_status gps_scene_upload_range(gps_scene* scene, uint64_t first, const gps_scene_desc* desc);
void gps_scene_destroy(gps_scene* scene);
/* Device bytes held by the scene. */
uint64_t gps_scene_device_bytes(const gps_scene* scene);

/* Scene box of GPS_FORMAT_QUANTIZED (GaussianSceneBounds,
 * src/core/gaussians/gaussian_types.cuh:123-170): the position box widened
 * by 0.001 on each side, the box of log(scale), and for higher SH
 * coefficient j (0..14) and channel c the range [min_sh[3j+c],
 * max_sh[3j+c]] (coefficients above sh_degree count as 0). */
typedef struct gps_quantized_bounds {
    float min_pos[3], max_pos[3];
    float min_log_scale[3], max_log_scale[3];
    float min_sh[45], max_sh[45];
} gps_quantized_bounds;
gps_status gps_quantized_bounds_compute(const gps_scene_desc* desc, gps_quantized_bounds* out);
/* GPS_FORMAT_QUANTIZED needs its box before gps_scene_upload_range
 * (gps_scene_create computes it from the whole description).  A box wider
 * than the data is allowed; positions outside it are clamped to it. */
gps_status gps_scene_set_quantized_bounds(gps_scene* scene, const gps_quantized_bounds* bounds);
/* Host encoder of GPS_FORMAT_QUANTIZED (the code gps_scene_upload_range
 * uses; encode_gaussian_geometry :237-320, normalized_float_to_byte :204-207,
 * CompactSH's constructor sh.cuh:29-39): geometry N*4 words, opacity N
 * bytes, sh_dc N*3 floats, sh_codes N*45 bytes (coefficient-major, rgb
 * interleaved).  A component whose box has zero width (0/0) encodes as 0. */
gps_status gps_encode_quantized(const gps_scene_desc* desc, const gps_quantized_bounds* bounds, uint32_t* geometry,
                                uint8_t* opacity, float* sh_dc, uint8_t* sh_codes);

/* ------------------------------------------------------------------ passes */
/* Replaces GaussianPointSplatting::GaussianPointSplatting(const RenderSettings&)
 * (gaussian_point_splatting.cu:777-816) for an already-loaded scene.  The scene
 * must outlive the pass; several passes may share one scene.
 * GPS_ERR_UNSUPPORTED when width or height times the supersampling factor
 * reaches 65536, or the framebuffer (64x64 tiles x samples per pixel) would
 * hold 2^31 words or more; the same limits apply per frame and per batch. */
gps_status gps_pass_create(const gps_render_settings* settings, gps_scene* scene, gps_pass** out);
void gps_pass_destroy(gps_pass* pass);
/* Splat kernel choice.  GPS_SPLAT_BINNED: screen regions rendered in LDS,
 * no per-sample global atomics (spp <= 4, at most 8192 regions per frame,
 * else point-parallel).  GPS_SPLAT_POINT_PARALLEL: one 64-bit global
 * atomicMin per accepted sample.  GPS_SPLAT_AUTO (default): binned when a
 * point has at least 8 samples (K * spp) and the pass has at most 8 samples
 * per framebuffer word (points * K <= 8 * upscaled width * height, tested on
 * the device after the point count; denser passes splat point-parallel and
 * report gps_stats.splat_fallback = 2).  All give the same framebuffer bit
 * for bit.  No reference counterpart (the reference has only the
 * point-parallel form, gaussian_point_splatting.cu:383-432). */
#define GPS_SPLAT_AUTO 0
#define GPS_SPLAT_POINT_PARALLEL 1
#define GPS_SPLAT_BINNED 2
gps_status gps_pass_set_splat_mode(gps_pass* pass, int32_t mode);
/* Diagnostic counters, off by default.  GPS_COUNT_ATOMICS: the splat's
 * global atomics into gps_stats.atomics (one atomic per workgroup on a single
 * word).  GPS_COUNT_REJECTS: per render pass, why each Gaussian of an
 * unculled block did not become valid, read with GPS_DBG_REJECTS (the
 * preprocess then always runs its work-list form, same outputs).  No
 * reference counterpart (the reference's statistics are the totals of
 * gaussian_point_splatting.cu:675-685). */
#define GPS_COUNT_ATOMICS 1
#define GPS_COUNT_REJECTS 2
gps_status gps_pass_set_counters(gps_pass* pass, int32_t flags);
/* Group occlusion test, on by default (enable = 1): in frames with occlusion
 * culling and a max-only Hi-Z chain (occlusion_culling_min_reduce_count 0), a
 * conservative test per group of 64 consecutive Gaussians against the pass's
 * Hi-Z chain lets the preprocess skip groups whose every member the
 * per-Gaussian Hi-Z test (gaussian_point_splatting.cu:226-242) would reject.
 * Outputs are the same bits with it on or off.  No reference counterpart (the
 * reference tests every Gaussian of an unculled block). */
gps_status gps_pass_set_group_culling(gps_pass* pass, int32_t enable);
/* Ordered point-parallel splat.  GPS_ORDERED_AUTO (default): a pass whose
 * points have at least 8 samples (K * spp) and which has more than 32
 * samples per framebuffer word (points * K > 32 * upscaled width * height,
 * tested on the device) is splatted nearest first: chunks of points
 * are bucketed by the depth of their first Gaussian and splatted in four
 * phases (the nearest 1/8, 1/4, 1/2 and all of the chunks); before each
 * later phase (and before the first one in a second render pass) the maximum
 * depth of every 16x16-pixel framebuffer tile is taken, and a Gaussian whose
 * bounds cover at most 16 tiles, all nearer than it, is skipped, since none
 * of its samples can lower a word.  GPS_ORDERED_OFF: chunk order, nothing
 * skipped.  GPS_ORDERED_ALWAYS: ordered whatever the density.  The
 * framebuffer is the same bit for bit in every mode; the unbiased splat
 * only.  No reference counterpart (the reference splats every point,
 * gaussian_point_splatting.cu:383-432). */
#define GPS_ORDERED_AUTO 0
#define GPS_ORDERED_OFF 1
#define GPS_ORDERED_ALWAYS 2
gps_status gps_pass_set_ordered_splat(gps_pass* pass, int32_t mode);
/* Device memory held by the pass (scene not included), and what
 * gps_pass_create would allocate for these settings. */
uint64_t gps_pass_device_bytes(const gps_pass* pass);
uint64_t gps_pass_estimate_bytes(const gps_render_settings* settings, const gps_scene* scene);

/* Replaces RenderPass::render(glm::vec3* d_image, float* d_depth, const
 * RenderSettings&) (src/core/rendering/render_pass.cuh:17-24) as implemented
 * by GaussianPointSplatting::render (gaussian_point_splatting.cu:599-737).
 * d_color: device float[W*H*3] average linear RGB; d_depth: device float[W*H]
 * minimum NDC depth.  Enqueued on `stream`; returns without waiting. */
gps_status gps_pass_render(gps_pass* pass, const gps_frame_params* frame,
                           float* d_color, float* d_depth, void* stream);

/* Statistics of the last frame; synchronizes `stream`. */
gps_status gps_pass_stats(gps_pass* pass, gps_stats* out, void* stream);

/* Test/inspection access to the pass's internal device state after a render,
 * copied to host memory (synchronizes `stream`).  `count` is in elements and
 * must not exceed gps_pass_debug_size(). */
typedef enum gps_debug_buffer {
    GPS_DBG_WEIGHTS = 0,        /* uint32[N]: point count of valid Gaussians, 0 otherwise (last pass) */
    GPS_DBG_MASK_CURRENT = 1,   /* uint32[ceil(N/32)] valid-mask of the last pass */
    GPS_DBG_MASK_PREVIOUS = 2,  /* uint32[ceil(N/32)] valid-mask of the pass before */
    GPS_DBG_VALID_IDS = 3,      /* uint32[V] ids of valid Gaussians, ascending */
    GPS_DBG_VALID_OFFSETS = 4,  /* uint32[V] exclusive point offsets of those ids */
    GPS_DBG_INDICES = 5,        /* uint32[P] Gaussian id of every point (materialized on request) */
    GPS_DBG_FRAMEBUFFER = 6,    /* uint64[F] framebuffer words of the current frame */
    GPS_DBG_RECORDS = 7,        /* 16 x uint32 per Gaussian: cached splat parameters */
    GPS_DBG_HIZ = 8,            /* float[mip chain] Hi-Z pyramid, coarsest level first */
    GPS_DBG_CULL_MASK = 9,      /* uint32[ceil(blocks/32)] culled-block mask of the last pass */
    GPS_DBG_REJECTS = 10        /* uint32[2][9] per pass (GPS_COUNT_REJECTS): Gaussians of unculled
                                   blocks tested; valid in the first pass (second pass only);
                                   opacity <= 0; projection rejected (near plane, NDC z, det <= 0);
                                   footprint off screen; small-eigenvalue cull; Hi-Z occluded;
                                   point count 0; valid; then uint32[2]: per pass, members of
                                   groups the group occlusion test marked that survived the
                                   per-Gaussian tests (0 when the test is sound); then
                                   uint32[2]: per pass, points the ordered splat skipped */
} gps_debug_buffer;
uint64_t gps_pass_debug_size(gps_pass* pass, int32_t which);
gps_status gps_pass_debug_copy(gps_pass* pass, int32_t which, void* host_dst, uint64_t count, void* stream);

/* Per-stage timing with HIP events recorded on the render stream (no host
 * synchronization while rendering).  capacity = number of frames whose events
 * are kept (ring); 0 disables.  The summary synchronizes, returns the summed
 * milliseconds and launch counts per stage over the frames recorded since the
 * last summary, and resets.  Stages: 0 clear, 1 block cull, 2 preprocess,
 * 3 distribute (tile scan + compaction), 4 splat, 5 combine, 6 Hi-Z. */
#define GPS_NUM_STAGES 7
gps_status gps_pass_set_timing(gps_pass* pass, int32_t capacity);
gps_status gps_pass_timing_summary(gps_pass* pass, double out_ms[GPS_NUM_STAGES],
                                   uint64_t out_launches[GPS_NUM_STAGES], uint64_t* out_frames);

/* One frame with no device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The frame
 * is what gps_pass_render gives on a pass created with this frame's size and
 * spp, fully_disable_hierarchical_culling = 1 and max_number_of_points =
 * max_points, rendering this one frame (bit for bit the same colour and
 * depth); every intermediate buffer (weights, valid masks, splat records,
 * tile sums, compacted entries, chunk starts, counters, framebuffer) lives in
 * a static device workspace of GPS_NOALLOC_SCRATCH_BYTES that the library
 * holds on every device it runs on until the process ends.  Limits: at most
 * GPS_NOALLOC_MAX_GAUSSIANS Gaussians, a framebuffer of at most
 * GPS_NOALLOC_MAX_FB_WORDS words (64x64 tiles x spp; 1920x1080 at ss 1 and
 * 1 spp fits), max_points <= GPS_NOALLOC_MAX_POINTS; frames with occlusion
 * or block culling are refused (both need state kept between frames).
 * Beyond a limit: GPS_ERR_UNSUPPORTED, nothing enqueued.  Every call on one
 * device shares its workspace, which keeps the last call's contents (each
 * call clears what must start at zero): calls that overlap on one device
 * race, so enqueue them on one stream or order them with events. */
#define GPS_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define GPS_NOALLOC_MAX_GAUSSIANS (1ull << 19)
#define GPS_NOALLOC_MAX_FB_WORDS (1ull << 21)
#define GPS_NOALLOC_MAX_POINTS (1ull << 26)
gps_status gps_frame_render_noalloc(gps_scene* scene, const gps_frame_params* frame, int64_t max_points,
                                    float* d_color, float* d_depth, void* stream);

/* ---------------------------------------------------------------- resolve */
typedef enum gps_accumulation_mode {
    GPS_ACCUM_PROGRESSIVE = 0,
    GPS_ACCUM_TAA = 1,
    GPS_ACCUM_NONE = 2
} gps_accumulation_mode;

/* TAA reprojection (post_processing_kernels.cuh:222-285, launched at
 * render_server.cu:126-142).  Reads cur_color (this frame's combine output)
 * and its 3x3 neighbourhood, reprojects into the history (prev_color /
 * prev_depth) with `reprojection` (column-major prevVP * affineInverse(view) *
 * inverse(proj)), blends 0.9 towards the clamped history, writes the blended
 * colour to out_color and RGB8 to out_rgb8.  frame_number > 0 enables the
 * blend.  out_color must not alias cur_color (the reference blends in place;
 * here the neighbourhood is always the un-blended frame). */
gps_status gps_resolve_taa(const float* cur_color, const float* cur_depth,
                           const float* prev_color, const float* prev_depth,
                           const float reprojection[16], float blend_factor,
                           int32_t width, int32_t height, int32_t frame_number,
                           float* out_color, uint8_t* out_rgb8, void* stream);

/* Progressive accumulation (post_processing_kernels.cuh:108-140, launched at
 * render_server.cu:143-156): color += accum (if num_frames > 1), accum = 0,
 * rgb8 = round(clamp(color * 255 / num_frames)). Updates `color` in place. */
gps_status gps_resolve_progressive(float* color, float* accum, int32_t width, int32_t height,
                                   int32_t num_frames, uint8_t* out_rgb8, void* stream);

/* ----------------------------------------------------------- camera math */
/* Host helpers reproducing the reference's camera conventions (glm 1.0.2
 * right-handed, NDC z in [-1, 1]). */
/* glm::lookAtRH (glm/ext/matrix_transform.inl:153-173). */
void gps_look_at(const float eye[3], const float center[3], const float up[3], float out_view[16]);
/* glm::perspective(radians(fovy), aspect, near, far) (src/world/camera.cpp:96-98). */
void gps_perspective(float fovy_degrees, float aspect, float near_plane, float far_plane, float out_proj[16]);
/* Camera::getFrustumPlanesFromViewMatrix (src/world/camera.cpp:126-155). */
void gps_frustum_planes(float fovy_degrees, const float view[16], float aspect, float near_plane,
                        float far_plane, float out_planes[24]);
/* Fills the camera-derived fields of `frame` (view, proj, planes, fovy,
 * near/far, width/height, ss) from a look-at camera, with the reference's
 * defaults for every flag (unbiased splatting on, point reduction on, both
 * culls on). */
void gps_frame_params_init(gps_frame_params* frame, const float eye[3], const float center[3],
                           const float up[3], float fovy_degrees, int32_t width, int32_t height,
                           int32_t supersampling_factor, float near_plane, float far_plane);
/* ------------------------------------------------------- camera paths */
/* One camera keyframe of a camera-path file (CameraLoader::CameraState,
 * src/utils/camera_loader.cuh:26-37; the files under assets/cameras): rotation is
 * the file's 3x3 camera-to-world rotation, row-major as written. */
typedef struct gps_camera_state {
    int32_t id;
    int32_t width, height;
    float position[3];
    float rotation[9];
    float fy, fx;
    char img_name[64];
} gps_camera_state;
/* Reads a camera-path JSON file (loadCameras, camera_loader.cu:120-148):
 * all cameras sorted by id, the first `capacity` copied to `out`, the total
 * in *count (capacity 0 only counts).  GPS_ERR_IO for a missing or
 * malformed file. */
gps_status gps_camera_path_load(const char* path, gps_camera_state* out, int32_t capacity, int32_t* count);
/* Writes cameras in the same JSON layout (saveCameras, camera_loader.cu:106-117). */
gps_status gps_camera_path_save(const char* path, const gps_camera_state* states, int32_t count);
/* Vertical field of view of a camera, degrees (Camera::focal2fovdeg(fy,
 * height), src/world/camera.h:44-47). */
float gps_camera_fovy(const gps_camera_state* state);
/* Resolution a camera file gives the renderer when none is forced
 * (fromCameraState, camera_loader.cu:73-85, load_cam_resolution_height_only):
 * the width is kept and the height follows the camera's aspect ratio. */
gps_status gps_camera_resolution(const gps_camera_state* state, int32_t width, int32_t* out_width,
                                 int32_t* out_height);
/* Frame of a camera state as the viewer builds it (fromCameraState + Transform
 * + main.cu:146-158): view = scale(1,-1,-1) * affineInverse(translate(position)
 * * mat4_cast(quat_cast(basis))), perspective projection, frustum planes;
 * fovy_degrees <= 0 takes gps_camera_fovy(state).  Flags as
 * gps_frame_params_init. */
gps_status gps_frame_params_from_camera(const gps_camera_state* state, int32_t width, int32_t height,
                                        float fovy_degrees, int32_t supersampling_factor, float near_plane,
                                        float far_plane, gps_frame_params* out);
/* Frame of camera-path playback at time_s seconds (startCameraPath +
 * CameraPath::interpolate, camera_loader.cu:93-104, src/world/camera.h:53-106):
 * keyframes 1/24 s apart, position by mix, rotation by slerp, fovy of the
 * first camera unless given; *finished = 1 at the end of the path. */
gps_status gps_camera_path_frame(const gps_camera_state* states, int32_t count, float time_s, float fovy_degrees,
                                 int32_t width, int32_t height, int32_t supersampling_factor, float near_plane,
                                 float far_plane, gps_frame_params* out, int32_t* finished);

/* deltaMatrix of the TAA pass (render_server.cu:128-129). */
void gps_taa_reprojection(const float prev_view_proj[16], const float view[16], const float proj[16],
                          float out[16]);
/* proj * view, the previousViewProjectionMatrix of the next frame (main.cu:176). */
void gps_mat4_mul(const float a[16], const float b[16], float out[16]);

/* -------------------------------------------------------- synthetic scenes */
/* Deterministic synthetic scenes (splitmix64 + FP64 Box-Muller):
 *   kind 0: uniform box   positions U([-half,half]^3), log-scale N(mu, sigma)
 *   kind 1: clustered     `clusters` centres U([-half,half]^3), members
 *                          N(centre, cluster_sigma^2)
 * Quaternions normalize(N(0,1)^4); opacity sigmoid(N(0, opacity_sigma));
 * SH DC N(0, dc_sigma), higher bands N(0, rest_sigma) up to sh_degree.
 * Arrays sized as in gps_scene_desc; index order = generation order. */
typedef struct gps_synth_params {
    uint64_t seed;
    uint64_t num_gaussians;
    int32_t kind;
    int32_t sh_degree;
    int32_t clusters;
    int32_t reserved;
    double half_extent;
    double cluster_sigma;
    double log_scale_mu, log_scale_sigma;
    double opacity_sigma;
    double dc_sigma, rest_sigma;
    double offset[3];   /* translation added to every position */
} gps_synth_params;
gps_status gps_synth_generate(const gps_synth_params* p, uint64_t first, uint64_t count, float* means,
                              float* rotations, float* scales, float* opacities, float* shs);
/* With rotations, scales, opacities and shs all NULL only the positions are
 * written (cheap pass for computing a load-time order of a large scene). */
/* Same Gaussians for an arbitrary index list: Gaussian `indices[k]` of the
 * scene into slot k (each Gaussian has its own stream, so order is free). */
gps_status gps_synth_generate_indices(const gps_synth_params* p, const uint64_t* indices, uint64_t count,
                                      float* means, float* rotations, float* scales, float* opacities, float* shs);

/* Load-time Morton ordering, the reference's --sort-morton-order
 * (src/core/gaussians/gaussian_loader.cu:363-426): writes perm[new] = old for
 * a stable sort of the 63-bit Morton codes of the bbox-normalized positions. */
gps_status gps_morton_order(const float* means, uint64_t n, uint32_t* perm);
/* Load-time BVH order, the reference's default for a scene that is neither
 * Morton-sorted nor marked "bvh_sorted_order" (GaussianBVH::
 * build_bvh_and_reorder, src/core/rendering/occlusion/gaussian_bvh.cu:146-169,
 * called from gaussian_point_splatting.cu:791-797): the tinybvh 1.6.7
 * binned-SAH build (packages/tinybvh/tiny_bvh.h:2292-2420) over the 3-sigma
 * boxes, the right-first leaf walk of collect_leaf_order (:14-34) and the
 * cycle reorder (:47-84).  Writes perm[new] = old.  Built on the current GPU
 * from host arrays laid out as in gps_scene_desc (rotations w, x, y, z;
 * activated scales).  GPS_ERR_UNSUPPORTED above 2^31 - 1 Gaussians, or for
 * scenes whose node areas exceed ~1e30 / count (where tinybvh splits on a
 * degenerate axis).  A load-time call: it allocates its own device memory
 * (about 80 bytes per Gaussian plus an upload buffer of at most 336 MB),
 * works on the null stream and returns after
 * the device has finished. */
gps_status gps_bvh_order(const float* means, const float* rotations, const float* scales, uint64_t n,
                         uint32_t* perm);

/* ---------------------------------------------------------- PLY files */
/* 3D Gaussian Splatting PLY (binary little-endian), as the reference's
 * GaussianLoader reads and writes it (src/core/gaussians/gaussian_loader.cu:
 * 105-427 load, 455-562 write).  Loading applies the reference's
 * activations: opacity clamp(sigmoid), scale exp, rotation normalized and
 * kept (w, x, y, z), SH DC + f_rest de-interleaved to coefficient-major RGB
 * (shs has 48 floats per Gaussian; all higher bands 0 unless f_rest_44 is
 * declared).  sort_morton_order reorders like --sort-morton-order unless the
 * file has the "bvh_sorted_order" comment. */
typedef struct gps_ply_info {
    uint64_t num_gaussians;
    uint64_t record_bytes;
    int32_t has_rest;      /* f_rest_0..44 declared */
    int32_t bvh_sorted;    /* "bvh_sorted_order" comment present */
} gps_ply_info;
gps_status gps_ply_info_read(const char* path, gps_ply_info* out);
gps_status gps_ply_load(const char* path, int32_t sort_morton_order, uint64_t capacity, float* means,
                        float* rotations, float* scales, float* opacities, float* shs, uint64_t* loaded);
/* Writes the reference's full record layout (zero normals, opacity logit,
 * log scale); mark_sorted adds the "bvh_sorted_order" comment. */
gps_status gps_ply_write(const char* path, const gps_scene_desc* desc, int32_t mark_sorted);

/* ----------------------------------------------------- compact encoding */
/* Host encoder of GPS_FORMAT_HALF_FP8 (the same code gps_scene_create uses):
 * fp16 RNE rotation (N*4), scale (N*3), opacity (N), SH DC (N*3); the
 * (deg+1)^2-1 higher SH coefficients per channel as FP8 E4M3 RNE codes of
 * c / sh_scale (N*45, coefficient-major, rgb interleaved, zero-padded) with
 * sh_scale = max|c| / 448 (1 if all are zero). */
gps_status gps_encode_half_fp8(const gps_scene_desc* desc, uint16_t* rotations, uint16_t* scales, uint16_t* opacities,
                               uint16_t* sh_dc, uint8_t* sh_rest, float* sh_scale);
uint16_t gps_float_to_half(float f);
uint8_t gps_float_to_fp8e4m3(float f);

/* ------------------------------------------------------------- batches */
/* Many camera views of one scene per launch sequence.  No reference
 * counterpart (the reference renders one camera per frame,
 * gaussian_point_splatting.cu:599-737).  Up to max_views frames of uneven
 * size / spp / ss go through one block-cull launch, one preprocess launch
 * (each Gaussian fetched once for all views), one scan + distribution over
 * the concatenated per-view tiles, one persistent splat over all views'
 * points, and a combine per view.  View k's colour / depth are bit-identical
 * to the same frame rendered alone by gps_pass_render; occlusion culling is
 * not available in a batch (it needs the previous frame of the same camera).
 * With use_graph the launches are captured once per view layout (sizes,
 * output pointers, stream) into a HIP graph and replayed for new cameras;
 * a graph needs a non-null stream. */
typedef struct gps_batch gps_batch;
typedef struct gps_batch_settings {
    int32_t max_views;                   /* 1..256 */
    int32_t max_width, max_height;
    int32_t max_spp;
    int32_t max_supersampling;
    int32_t fully_disable_hierarchical_culling;
    int32_t gaussians_per_block;         /* 0 = max(4, floor(N / 1e6)) */
    int32_t reserved;
    int64_t max_number_of_points;        /* point budget per view */
} gps_batch_settings;
typedef struct gps_batch_tuning {
    int32_t splat_threads;               /* workgroup size of the splat */
    int32_t splat_points_per_thread;     /* points per thread (x 64 = points per wave) */
    int32_t use_graph;
    int32_t variants;                    /* filled by gps_batch_autotune */
    int32_t variant_threads[8];
    int32_t variant_points_per_thread[8];
    float variant_ms[8];                 /* batch time of each variant */
    int32_t preprocess_views_per_workgroup; /* views one preprocess workgroup evaluates (grid rows split
                                               the batch's views; 0: all of them) */
    int32_t pre_variants;                /* filled by gps_batch_autotune */
    int32_t pre_variant_views[4];
    float pre_variant_ms[4];             /* batch time of each split, with the chosen splat variant */
} gps_batch_tuning;
gps_status gps_batch_create(const gps_batch_settings* settings, gps_scene* scene, gps_batch** out);
gps_status gps_batch_render(gps_batch* batch, int32_t num_views, const gps_frame_params* views,
                            float* const* d_colors, float* const* d_depths, void* stream);
/* Times every splat variant on this batch (`repeats` batches each), then
 * every preprocess split (views per workgroup 0 = all, 16, 4) with the
 * fastest splat variant, keeps the fastest and reports all; the graph
 * setting is kept.  Outputs are the same bits in every variant. */
gps_status gps_batch_autotune(gps_batch* batch, int32_t num_views, const gps_frame_params* views,
                              float* const* d_colors, float* const* d_depths, void* stream, int32_t repeats,
                              gps_batch_tuning* out);
gps_status gps_batch_set_tuning(gps_batch* batch, const gps_batch_tuning* tuning);
/* Splat form per view, with the rule and modes of gps_pass_set_splat_mode:
 * the binned views of a batch share one set of binning launches (regions of
 * all of them in one job queue), the rest use the point-parallel batch
 * splat.  The binned buffers (16 B per compacted entry, 12 B per list slot,
 * about 1.5 slots per entry) are allocated on the first batch that bins a
 * view; if that allocation fails every view stays point-parallel.  No
 * reference counterpart. */
gps_status gps_batch_set_splat_mode(gps_batch* batch, int32_t mode);
/* Counters of view `view` of the last batch (atomics, bin_entries, bin_jobs,
 * bin_big: whole batch). */
gps_status gps_batch_view_stats(gps_batch* batch, int32_t view, gps_stats* out, void* stream);
uint64_t gps_batch_device_bytes(const gps_batch* batch);
void gps_batch_destroy(gps_batch* batch);

/* ------------------------------------------------------------- serving */
/* Serving loop with continuous batching over resident scenes.  No reference
 * counterpart: the reference renders one camera per frame from a single
 * host loop (RenderServer::renderFrame, render_server.cu:86-160).  Requests
 * are independent frames (camera, size <= max, spp <= max, RNG frame) of one
 * of the added scenes; occlusion culling is off for them (it needs the
 * previous frame of the same camera), block frustum culling stays on, and
 * the output is the progressive resolve of that one frame.  A scheduler
 * thread admits every arrived request that finds an idle slot (HIP stream)
 * with a pass for its scene, or HBM budget to create one (evicting idle
 * passes least-recently-used); work of different slots overlaps on the
 * GPU.  Results are bit-identical whatever ran beside them. */
typedef struct gps_server gps_server;
typedef struct gps_server_config {
    int32_t slots;                 /* requests in flight on the GPU, 1..64 */
    int32_t max_width, max_height; /* largest request */
    int32_t max_spp;
    int32_t max_supersampling;
    int32_t tick_us;               /* scheduler period */
    int32_t device;
    uint64_t hbm_budget_bytes;     /* for slot buffers + passes; 0 = 90% of free HBM */
} gps_server_config;
typedef struct gps_request {
    int32_t scene;                 /* id from gps_server_add_scene */
    int32_t reserved;
    gps_frame_params frame;        /* width/height/samples_per_pixel/frame per request */
    uint8_t* out_rgb8;             /* host, width*height*3, written on completion (may be NULL) */
    float* out_color;              /* host, width*height*3 floats (may be NULL) */
    double arrival_s;              /* server clock time the request arrives; < 0 = on submit */
} gps_request;
typedef struct gps_request_result {
    uint64_t ticket;
    int32_t status;                /* gps_status of the request */
    int32_t slot;
    double arrival_s, admit_s, done_s;  /* server clock, seconds */
} gps_request_result;
typedef struct gps_server_stats {
    uint64_t submitted, completed;
    uint64_t passes_created, passes_evicted;
    uint64_t max_queue, max_batch, ticks_with_admissions;
    uint64_t hbm_budget_bytes, hbm_used_bytes;
    double clock_s, slot_busy_s;
} gps_server_stats;
gps_status gps_server_create(const gps_server_config* config, gps_server** out);
/* The scene must outlive the server. */
gps_status gps_server_add_scene(gps_server* server, gps_scene* scene, int32_t* scene_id);
/* Optional, before start: create the pass of every (slot, scene) pair that
 * fits the budget, so no request pays for an allocation. */
gps_status gps_server_prepare(gps_server* server);
gps_status gps_server_start(gps_server* server);
double gps_server_clock(const gps_server* server);
gps_status gps_server_submit(gps_server* server, const gps_request* request, uint64_t* ticket);
/* Completed requests (up to max), waiting up to wait_ms for the first. */
int32_t gps_server_poll(gps_server* server, gps_request_result* out, int32_t max, int32_t wait_ms);
gps_status gps_server_drain(gps_server* server, int32_t timeout_ms);
gps_status gps_server_get_stats(gps_server* server, gps_server_stats* out);
const char* gps_server_last_error(const gps_server* server);
void gps_server_destroy(gps_server* server);

/* ---------------------------------------------------------- self-tests */
/* Device evaluation of the shared deterministic scalar math and RNG, used to
 * prove host/device bit-exactness.  No reference counterpart (diagnostics).
 * kind: 0 exp, 1 log, 2 sin, 3 cos, 4 sqrt, 5 reciprocal, 6 dilog,
 * 7 inv_dilog, 8 float->int round-to-nearest-even (saturating),
 * 9 Box-Muller x of (in[i], in[i+1]), 10 corrected Box-Muller y (alpha 0.7),
 * 11 FP8 E4M3 decode of code in[i], 12 the log's quotient x / (2 + x),
 * 13 the square root of the sampling transform (x = 0 or normal). */
gps_status gps_selftest_math(int32_t kind, const float* d_in, float* d_out, uint64_t n, void* stream);
/* Per item i: make_seed(i, frame) then `draws` pcg2d states (2+2*draws u32). */
gps_status gps_selftest_rng(uint32_t n, uint32_t frame, int32_t draws, uint32_t* d_out, void* stream);
/* Per item i: poisson(make_seed(i, 12345), lambda[i]). */
gps_status gps_selftest_poisson(uint32_t n, const float* d_lambda, uint32_t* d_out, void* stream);
/* Streaming device copy (bytes % 16 == 0): reference kernel for HBM byte counters. */
gps_status gps_selftest_copy(void* d_dst, const void* d_src, uint64_t bytes, void* stream);
/* blocks x 256 threads each issue `iters` 64-bit atomicMin at hashed random
 * words of d_buf[0, words) (words < 2^32): the random framebuffer-update rate
 * the point-parallel splat's atomics are priced against. */
gps_status gps_selftest_atomics(uint64_t* d_buf, uint64_t words, uint32_t blocks, uint32_t iters, void* stream);

/* ------------------------------------------------------- device plumbing */
/* Minimal HIP plumbing for hosts without a HIP binding (ctypes, cgo, JNI). */
gps_status gps_device_count(int32_t* out);
gps_status gps_set_device(int32_t device);
gps_status gps_device_malloc(void** out, uint64_t bytes);
gps_status gps_device_free(void* ptr);
gps_status gps_memcpy_h2d(void* dst, const void* src, uint64_t bytes, void* stream);
gps_status gps_memcpy_d2h(void* dst, const void* src, uint64_t bytes, void* stream);
gps_status gps_memset_d8(void* dst, int32_t value, uint64_t bytes, void* stream);
gps_status gps_stream_create(void** out);
gps_status gps_stream_destroy(void* stream);
gps_status gps_stream_synchronize(void* stream);
gps_status gps_device_free_memory(uint64_t* free_bytes, uint64_t* total_bytes);
/* Event-based timing on a stream. */
gps_status gps_event_create(void** out);
gps_status gps_event_destroy(void* ev);
gps_status gps_event_record(void* ev, void* stream);
gps_status gps_event_elapsed_ms(void* start, void* stop, float* out_ms);

#ifdef __cplusplus
}
#endif

#endif /* GPS_H */
