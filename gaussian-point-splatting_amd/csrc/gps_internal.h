// gps_internal.h — device-side data layout shared by the kernels and the host
// orchestration of the MI355X path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gps.h"
#include "gps_hiz.h"
#include "gps_project.h"

namespace gps {

// Resident scene, structure-of-arrays ("planes") so every attribute load of a
// wave is one contiguous 256-byte segment.
//
// GPS_FORMAT_FP32: planes of `stride` floats:
//   0..2 mean xyz, 3..6 quaternion wxyz, 7..9 scale xyz, 10 opacity,
//   11 + 3*i + c  SH coefficient i of channel c (i < (deg+1)^2)
// GPS_FORMAT_HALF_FP8:
//   means: 3 float planes
//   geo:   uint4 per Gaussian = 8 halves: qw qx qy qz sx sy sz opacity
//   shrec: 64 bytes per Gaussian: 3 halves DC (rgb), 1 pad half, float scale,
//          45 fp8-E4M3 rest coefficients (coefficient-major, rgb interleaved),
//          3 pad bytes
// GPS_FORMAT_QUANTIZED (the reference's compressed scene):
//   geo:     uint4 per Gaussian = the 4 quantized geometry words
//   qop:     one opacity byte per Gaussian
//   shrec:   64 bytes per Gaussian: DC rgb fp32, 45 coefficient bytes, 7 pad
//   qbounds: 102 floats = gps_quantized_bounds (scene box)
struct DevScene {
    const float* planes;        // FP32 format
    const float* means;         // HALF_FP8 format
    const uint4* geo;           // HALF_FP8 and QUANTIZED formats
    const uint4* shrec;         // HALF_FP8 and QUANTIZED formats (4 x uint4 per Gaussian)
    const uint8_t* qop;         // QUANTIZED format
    const float* qbounds;       // QUANTIZED format
    uint64_t stride;            // plane stride in elements
    uint32_t N;
    int32_t sh_degree;
    int32_t format;
};

struct Mat4 {
    float m[16];
};

// Cached per-Gaussian splat parameters written by preprocess for every valid
// Gaussian and read by the splat kernel (replaces the reference's per-point
// re-projection, gaussian_point_splatting.cu:416-423; the values are the same
// bits either way).  64 bytes = 4 x dwordx4.
struct SplatRecord {
    float pmx, pmy;            // pixel-space mean
    float c0, c1, c2;          // Cholesky factor of cov2d
    float conx, cony, conz;    // conic
    float opacity;
    float inv_alpha;           // 1 / max(opacity, 1e-37)
    float dilog_alpha;         // dilog(opacity)
    uint32_t bmin;             // x | y << 16 (tile-aligned bounds)
    uint32_t bmax;
    uint32_t packed_lo;        // depth28 << 36 | colour36
    uint32_t packed_hi;
    float alpha_t;             // alpha_threshold(opacity): the alpha test's exponent bound
};
static_assert(sizeof(SplatRecord) == 64, "record must stay 64 bytes");
static_assert(sizeof(gps_quantized_bounds) == 102 * sizeof(float), "scene box is read as 102 floats");


// Flat culling blocks (gaussian_bvh.cuh:28-53).
struct CullNode {
    float mn[3];
    float mx[3];
    uint32_t begin, count;
};
struct CullView {
    const CullNode* nodes;
    uint32_t* mask;
    uint32_t node_count;
    uint32_t mask_words;
    uint32_t gpb;   // Gaussians per block
};

// Group occlusion test (gps_hiz.h): group bounds built once per pass, one
// byte per group and pass.  Only groups with a member in an unculled block
// are written in a pass; the preprocess reads no other (their members are
// not tested).
struct GroupView {
    const GroupBox* boxes;
    uint8_t* occluded;    // 1: every member is rejected by the Hi-Z test (this pass)
    uint32_t count;       // ceil(N / GPS_GROUP); the byte array is padded to a multiple of 4
};

// Device-side counters of one render pass.
struct PassCounters {
    uint32_t total;       // sum of masked weights (mod 2^32), = reference last_offset + last_count
    uint32_t valid;       // number of valid Gaussians
    uint32_t points;      // min(total, budget)
    uint32_t chunks;      // ceil(points / CHUNK)
};
// Per frame: counters of pass 1 and 2, culled block counts, Hi-Z ticket.
struct FrameCounters {
    PassCounters pass[2];
    uint32_t culled_blocks[2];
    uint32_t hiz_ticket;
    uint32_t iteration;   // valid-mask ping-pong index (flips when a pass yields points)
    uint32_t atomics;     // 64-bit atomicMin issued by the splat kernel
    uint32_t culled_groups[2];  // per pass: groups whose members all skip the per-Gaussian steps
    uint32_t live_blocks[2];    // per pass: unculled blocks listed by the block cull for the group test
    uint32_t rejects[2][9];  // per pass: preprocess reject reasons (diagnostic; see RejectReason)
    uint32_t group_violations[2];  // per pass, with the reasons: members of culled groups that survived (must be 0)
    uint32_t ordered_skipped[2];   // per pass, with the reasons: points the ordered splat skipped (diagnostic)
};
static_assert(sizeof(FrameCounters) == 156, "counter words zeroed by the clear kernel");

// Region geometry of the binned splat for one framebuffer.
struct BinGeom {
    int32_t rw_shift, rh_shift;  // log2 of the region width / height in pixels
    int32_t nrx, nry;            // regions per row / column
    int32_t up_w, up_h, spp;
    int32_t tiles_per_row;       // framebuffer 64x64 tiles per row
    int32_t rtx, rty;            // framebuffer tiles per region (x, y)
    uint32_t nregions;
};

// One view of a multi-view batch (device table entry).  The host fills the
// frame part before every batch; the device writes the per-view results.
struct BatchView {
    FrameConsts f;          // f.frame_seed = 4 * frame (preprocess); the splat uses + 1
    float planes[24];       // frustum planes for block culling
    uint64_t fb_offset;     // first framebuffer word of this view
    uint64_t fb_words;      // words of this view's framebuffer
    uint64_t background;    // background word
    uint32_t budget;        // point budget of the view
    BinGeom bg;             // binned splat regions (bg.nregions == 0: point-parallel splat)
    uint32_t rbase;         // first region of the view in the batch's region numbering
    uint32_t dense_points;  // more points than this: the view is splatted point-parallel
    // written on the device
    uint32_t base;          // first point of the view in the batch's point list
    uint32_t points;        // points splatted (min(total, budget))
    uint32_t total;         // uncapped sum of point counts
    uint32_t valid;         // valid Gaussians
    uint32_t ebase;         // first compacted entry of the view
    uint32_t culled_blocks;
};
#define GPS_BATCH_MAX_VIEWS 256

// Splat work decomposition: one workgroup processes CHUNK consecutive points,
// PPT consecutive points per thread.
#ifndef GPS_SPLAT_THREADS
#define GPS_SPLAT_THREADS 256
#endif
#ifndef GPS_SPLAT_PPT
#define GPS_SPLAT_PPT 2
#endif
#define GPS_SPLAT_CHUNK (GPS_SPLAT_THREADS * GPS_SPLAT_PPT)
#ifndef GPS_SPLAT_CHUNK_SHIFT
#define GPS_SPLAT_CHUNK_SHIFT 9u  // log2(GPS_SPLAT_CHUNK)
#endif
static_assert((1u << GPS_SPLAT_CHUNK_SHIFT) == GPS_SPLAT_CHUNK, "chunk is a power of two");
// Point-parallel splat: samples of one chunk that land on the same
// framebuffer word merge in a direct-mapped LDS table of this many words
// (0 = off), flushed once per chunk; workgroups per CU of the persistent grid.
// Measured over the 100M fly-through (splat ms per frame): no table 1.30;
// 1024 words 1.22; 2048 words 0.997 at 5 workgroups per CU (the LDS limit,
// 28.7 KB each) and 0.986 at 10 (oversubscribed); 4096 words 1.09.
#ifndef GPS_SPLAT_SLOTS
#define GPS_SPLAT_SLOTS 2048
#endif
#ifndef GPS_SPLAT_WGS_PER_CU
#define GPS_SPLAT_WGS_PER_CU 10
#endif
// A sample whose slot holds another word tries the neighbouring slot before
// the framebuffer (0.986 -> 0.979 ms; 16 workgroups per CU: 1.012).  Chunk
// size against the 2048-word table (splat ms per frame): 256 threads x 4
// points 0.972, x 2 points 0.913 (the table overflows less), x 1 point 0.940
// (the distributor's chunk starts cost more); 512 x 4 1.107, 256 x 8 1.175,
// 1024 x 2 with 4096 words 1.058; 256 x 2 with 1024 words 1.043; 512 x 1
// (eight waves per table) 0.962.
#ifndef GPS_SPLAT_PROBE2
#define GPS_SPLAT_PROBE2 1
#endif
// ...and then slot ^ 2 (512-point chunks: 0.913 -> 0.908 ms).
#ifndef GPS_SPLAT_PROBE3
#define GPS_SPLAT_PROBE3 1
#endif
static_assert((GPS_SPLAT_SLOTS & (GPS_SPLAT_SLOTS - 1)) == 0, "slots: a power of two");
// Binned splat: one 1024-thread workgroup renders a screen region of at most
// GPS_BIN_WORDS framebuffer words in LDS; GPS_BIN_PPT points per thread per
// step; Gaussians overlapping more regions than GPS_BIN_MAX_REGIONS take the
// global-atomic path; at most GPS_BIN_HIST regions per frame.
#define GPS_TILE_WORDS 4096u
#define GPS_BIN_THREADS 1024
#define GPS_BIN_PPT 4
#define GPS_BIN_CHUNK (GPS_BIN_THREADS * GPS_BIN_PPT)
#define GPS_BIN_WORDS 16384u
#define GPS_BIN_MAX_REGIONS 4u
#define GPS_BIN_HIST 8192u
#define GPS_BIN_JOB_POINTS 16384u
// GPS_SPLAT_AUTO bins a frame when a point has >= GPS_BIN_AUTO_SAMPLES
// samples (K * spp); with fewer the binning launches and the replays of
// Gaussians that span several regions cost more than the global atomics
// they save (over the whole 100M fly-through at 1 sample per point: 2.42 ms
// binned against 1.40 ms point-parallel per frame).
#define GPS_BIN_AUTO_SAMPLES 8
// ... and while the pass has at most GPS_BIN_DENSE_SAMPLES samples per
// framebuffer word (points * K <= 8 * up_w * up_h, tested on the device once
// the point count is known).  Denser frames reject most samples at the
// point-parallel splat's pre-read, which the binned form cannot skip:
// 1920x1080 ss 2, 4 spp, K 4 requests measured 0.5-1.05 ms binned vs
// 0.6-1.9 ms point-parallel at 2.4-4.8 samples per word, 2.44 vs 2.70 at
// 7.5, 2.17 vs 1.6 at 15, 14.6 vs 9.1 at 250.
#define GPS_BIN_DENSE_SAMPLES 8u

// Ordered point-parallel splat, for dense passes of points with many samples
// (K * spp >= GPS_BIN_AUTO_SAMPLES, and more than GPS_ORD_DENSE_SAMPLES
// samples per framebuffer word, tested on the device).  Chunks are put in log-depth buckets by the depth of their first Gaussian
// (GPS_ORD_BINS buckets: depths below 128 one each, then 128 steps per power
// of two) and splatted nearest first in GPS_ORD_PHASES phases holding the
// first 1/8, 1/4, 1/2 and all of the chunks.  Before every later phase the
// maximum depth of each 16x16-pixel framebuffer tile (all sample layers) is
// taken; a Gaussian whose bounds (16-pixel aligned) cover at most
// GPS_ORD_MAX_TILES tiles, all with a maximum depth below its own, is skipped:
// words only decrease, so every one of its samples would lose the min.
// Chunks of a phase go to the XCDs in groups of GPS_ORD_GROUP consecutive
// chunks (screen-coherent runs stay in one L2, the phase is spread over all).
#define GPS_ORD_BINS 2816u
#ifndef GPS_ORD_PHASES
#define GPS_ORD_PHASES 4
#endif
#ifndef GPS_ORD_GROUP
#define GPS_ORD_GROUP 16u
#endif
#define GPS_ORD_MAX_TILES 16
#define GPS_ORD_MIN_CHUNKS 64u  // fewer chunks: one phase in chunk order
// Measured on whole-scene requests at 1920x1080 / 1280x720 / 854x480, ss 2
// (ms per request, plain -> ordered): 100M Gaussians at 4 spp (55-250
// samples per word) 14.3 -> 10.4 with 55 % of the points skipped; at 1 spp
// (4 samples per point) 6.56 -> 6.48 although 66 % were skipped (the record
// loads, the chunk bookkeeping and the tile passes stay); 10M Gaussians
// (7.5-31 samples per word) slower by 0.13.  Occlusion-culled passes skip
// next to nothing (the Hi-Z test already removed those Gaussians).
#define GPS_ORD_DENSE_SAMPLES 32u

// Distributor tile: 256 threads x 16 Gaussians.
#define GPS_TILE_SHIFT 12
#define GPS_TILE (1u << GPS_TILE_SHIFT)
#ifndef GPS_PRE_THREADS
#define GPS_PRE_THREADS 256
#endif
static_assert(GPS_PRE_THREADS % 64 == 0 && (GPS_TILE % GPS_PRE_THREADS) == 0, "whole waves, whole tiles");

#if defined(__HIP_DEVICE_COMPILE__)
// Byte `sel` (a constant 0..3) of `word` as an OCP FP8 E4M3 value, converted
// by the hardware (v_cvt_f32_fp8; exact for every code).
__device__ __forceinline__ float fp8_e4m3_to_float_hw(uint32_t word, int sel) {
    switch (sel) {
        case 0: return __builtin_amdgcn_cvt_f32_fp8((int)word, 0);
        case 1: return __builtin_amdgcn_cvt_f32_fp8((int)word, 1);
        case 2: return __builtin_amdgcn_cvt_f32_fp8((int)word, 2);
        default: return __builtin_amdgcn_cvt_f32_fp8((int)word, 3);
    }
}
#elif defined(__HIP__)
__device__ __forceinline__ float fp8_e4m3_to_float_hw(uint32_t, int) { return 0.0f; }
#endif

// Frame constants of gps_pass_render for a frame with `spp` samples per pixel
// (gaussian_point_splatting.cu:605-615).
void make_frame_consts(const gps_frame_params* fp, int spp, FrameConsts* f);

}  // namespace gps

// Resident scene behind the gps_scene handle.
struct gps_scene {
    gps::DevScene view{};
    void* planes = nullptr;  // FP32 planes or HALF_FP8 means
    void* geo = nullptr;
    void* shrec = nullptr;
    void* qop = nullptr;      // QUANTIZED opacity bytes
    void* qbounds = nullptr;  // QUANTIZED scene box (device copy)
    gps_quantized_bounds bounds{};
    bool have_bounds = false;
    uint64_t bytes = 0;
};
