// gps_kernels.h — launch interface of the gfx950 kernels (gps_kernels.hip).
#pragma once

#include "gps_internal.h"

namespace gps {

struct PreprocessArgs {
    DevScene scene;
    FrameConsts f;
    uint32_t* maskA;
    uint32_t* maskB;
    const uint32_t* iteration;
    uint32_t* weights;
    SplatRecord* recs;
    uint32_t* tile_sum;
    uint32_t* tile_cnt;
    CullView cull;
    HiZView hiz;
    const uint8_t* gocc = nullptr;  // occluded-group bytes of this pass (group cull), or null
    // one byte per workgroup, 0 when the block cull culled its block (and
    // zeroed its mask words), or null; needs gaussians_per_block a multiple of 256
    const uint8_t* wg_live = nullptr;
    uint32_t* rejects = nullptr;  // 2 x 9 reject-reason counters + 2 group checks (diagnostic), or null
};

struct DistributeArgs {
    uint32_t* tile_sum;
    uint32_t* tile_cnt;
    uint32_t ntiles;
    uint32_t* tile_off;
    uint32_t* tile_voff;
    PassCounters* ctr;
    uint32_t budget;
    const uint32_t* weights;
    uint32_t* cidx;
    uint32_t* coff;
    uint32_t* chunk_start;
    uint32_t* iteration;
    CullView cull;
    int32_t hier;
    uint32_t N;
};

struct SplatArgs {
    const SplatRecord* recs;
    const uint32_t* cidx;
    const uint32_t* coff;
    const uint32_t* chunk_start;
    const PassCounters* ctr;
    uint64_t* fb;
    uint32_t* atomics_count;
    int32_t spp, K, up_w, unbiased;
    uint32_t frame_seed;
    uint32_t grid;   // multiple of xcds
    int32_t xcds;    // accelerator complex dies (workgroups round-robin over them)
    const uint32_t* only_if;  // when set, the splat runs only if *only_if != 0
};

// Ordered point-parallel splat (dense passes, GPS_ORD_BINS): the splat
// arguments plus the ordering buffers.
struct OrderedArgs {
    SplatArgs s;
    int32_t up_w, up_h;       // framebuffer pixels of this frame
    int32_t forced;           // order whatever the density
    int32_t pre_test;         // the framebuffer holds an earlier pass: phase 0 tests too
    uint16_t* bucket;         // per chunk
    uint32_t* hist;           // GPS_ORD_BINS, zero between passes
    uint32_t* lim;            // GPS_ORD_PHASES
    uint32_t* tmax;           // per 16x16 tile of the framebuffer
    uint32_t ntiles;          // 16x16 tiles of this frame's framebuffer
    int32_t tw64;             // 64x64 tiles per framebuffer row
    uint32_t cus;
    uint32_t* skipped_count;  // points skipped (diagnostic; may be null)
};
void launch_ordered_splat(const OrderedArgs& a, hipStream_t st);

// Splat constants shared by the point-parallel and the binned splat.
struct SplatConsts {
    int32_t spp, K, up_w, unbiased;
    uint32_t frame_seed;
    int32_t xcds;
};

struct BinnedArgs {
    const SplatRecord* recs;
    const uint32_t* cidx;
    const uint32_t* coff;
    const PassCounters* ctr;
    BinGeom g;
    SplatConsts sc;
    uint64_t background;
    uint64_t* fb;
    uint2* erect;
    uint32_t* ew;
    uint32_t *reg_cnt, *reg_pts, *reg_ecnt, *reg_ptot, *reg_eoff, *reg_poff, *reg_cursor, *reg_jobs;
    uint32_t *job_reg, *job_p0;
    uint32_t *list, *lw, *lpre;
    uint32_t *tsum, *toff;
    uint32_t* big;
    uint32_t* state;      // 8 words: list length, jobs, big count, stand-down, job cursor, big cursor
    uint32_t list_cap, job_cap, job_points;
    uint32_t cus;
    uint32_t dense_points;  // more points than this: the binned kernels stand down (state[3] = 2)
    uint32_t* atomics_count;
};
bool binned_geometry(int up_w, int up_h, int spp, BinGeom* g);
// Points above which a frame of up_w x up_h pixels has more than
// GPS_BIN_DENSE_SAMPLES samples per framebuffer word (a point writes K * spp
// samples into up_w * up_h * spp words).
uint32_t binned_dense_points(int up_w, int up_h, int K);
void launch_binned_splat(const BinnedArgs& a, hipStream_t st);

// Binned splat buffers of a batch: entries are the batch's compacted entries,
// regions those of all binned views (view v owns [rbase, rbase + nregions)).
struct BatchBinArgs {
    int32_t any;                // at least one view is binned
    uint32_t nregions;          // regions of all binned views
    uint32_t max_view_regions;  // largest view's regions (LDS histogram size)
    uint2* erect;
    uint32_t* ew;
    uint32_t* big;
    uint32_t *reg_cnt, *reg_pts, *reg_ecnt, *reg_ptot, *reg_eoff, *reg_poff, *reg_cursor, *reg_jobs;
    uint32_t *job_reg, *job_p0;
    uint32_t *list, *lw, *lpre;
    uint32_t *tsum, *toff;
    uint32_t* state;            // as BinnedArgs::state
    uint32_t list_cap, job_cap, job_points;
    uint32_t cus;
};

// Multi-view batch (see the batch kernels): all arrays are concatenated per view.
struct BatchArgs {
    DevScene scene;
    BatchView* views;        // device table, nviews entries
    int32_t nviews;
    uint32_t npad;           // per-view stride of weights / records (multiple of the tile)
    uint32_t tiles_per_view;
    uint32_t* weights;
    SplatRecord* recs;
    uint32_t *tile_sum, *tile_cnt, *tile_off, *tile_voff;
    uint32_t *cidx, *coff;
    uint8_t* cview;
    uint32_t* chunk_start;
    PassCounters* ctr;
    uint32_t* iteration;     // scratch word (the scan flips it)
    CullView cull;           // mask = first view's mask
    uint32_t mask_stride;
    int32_t hier;
    uint64_t* fb;            // pool; views address it by fb_offset
    uint64_t max_fb_words;   // largest view
    int32_t fb_key32;        // every word of the batch's views has a 32-bit index (below 2^32 - 1)
    uint32_t grid;
    int32_t xcds;
    int32_t threads, ppt;    // splat variant
    int32_t pre_views;       // views one preprocess workgroup evaluates (0: all)
    uint32_t* atomics;
    BatchBinArgs bin;
};
extern const int kBatchSplatVariants[][2];
extern const int kNumBatchSplatVariants;
extern const int kBatchPreViews[];
extern const int kNumBatchPreViews;
void launch_batch_clear(const BatchArgs& a, hipStream_t st);
void launch_batch_cull(const BatchArgs& a, hipStream_t st);
void launch_batch_preprocess(const BatchArgs& a, hipStream_t st);
void launch_batch_distribute(const BatchArgs& a, hipStream_t st);
bool launch_batch_splat(const BatchArgs& a, hipStream_t st);
void launch_batch_binned(const BatchArgs& a, hipStream_t st);
void launch_batch_bin_big(const BatchArgs& a, hipStream_t st);

// Fills the framebuffer with `value`; also zeroes n0 words at z0 and n1 at z1.
void launch_clear(uint64_t* fb, uint64_t words, uint64_t value, hipStream_t st, uint32_t* z0 = nullptr,
                  uint32_t n0 = 0, uint32_t* z1 = nullptr, uint32_t n1 = 0);
// `live` (or null): the unculled blocks are listed there, their count in
// *live_count (zeroed by the frame's clear).  `wg_live` (or null): one byte
// per 256-Gaussian preprocess workgroup, 0 when its block is culled (blocks
// must be whole workgroups), and the culled blocks' valid-mask words zeroed
// in the preprocess's current mask (maskA / maskB by *iteration).
void launch_cull(const CullView& cv, const float planes[24], const Mat4& vp, const HiZView& h, int occl,
                 uint32_t* culled_count, uint32_t* live, uint32_t* live_count, uint32_t* maskA, uint32_t* maskB,
                 const uint32_t* iteration, uint8_t* wg_live, uint32_t N, hipStream_t st);
void launch_preprocess(const PreprocessArgs& a, hipStream_t st);
void launch_distribute(const DistributeArgs& a, hipStream_t st);
void launch_splat(const SplatArgs& a, hipStream_t st);
void launch_combine(const uint64_t* fb, float* color, float* depth, int W, int H, int ss, int spp, float near_plane,
                    float far_plane, int depth_only, hipStream_t st);
void launch_hiz_build(const HiZView& h, const float* depth, int kmin, uint32_t* ticket, hipStream_t st);
void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t st);
void launch_fill_f32(float* p, uint32_t n, float v, hipStream_t st);
void launch_build_nodes(const DevScene& s, uint32_t gpb, uint32_t nb, CullNode* nodes, hipStream_t st);
// Bounds of every group of GPS_GROUP Gaussians (once per pass), and the
// per-pass conservative group occlusion test against the pass's Hi-Z chain.
void launch_group_boxes(const DevScene& s, GroupBox* boxes, uint32_t count, hipStream_t st);
void launch_group_cull(const GroupView& gv, const FrameConsts& f, const HiZView& h, uint32_t* culled_count,
                       hipStream_t st);
void launch_group_cull_live(const GroupView& gv, const CullView& cv, const uint32_t* live, const uint32_t* live_count,
                            uint32_t N, const FrameConsts& f, const HiZView& h, uint32_t grid,
                            uint32_t* culled_count, hipStream_t st);
void launch_taa(const float* cur, const float* cur_depth, const float* prev, const Mat4& M, float blend, int W, int H,
                int frame_number, float* out, uint8_t* rgb8, hipStream_t st);
void launch_progressive(float* color, float* accum, int W, int H, int n, uint8_t* rgb8, hipStream_t st);
void launch_materialize_indices(const uint32_t* cidx, const uint32_t* coff, const PassCounters* ctr, uint32_t* out,
                                uint32_t max_points, hipStream_t st);

}  // namespace gps
