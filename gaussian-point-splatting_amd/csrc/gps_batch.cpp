// gps_batch.cpp — many camera views of one scene per launch sequence.
//
// No reference counterpart: the reference renders one camera per frame
// (GaussianPointSplatting::render, gaussian_point_splatting.cu:599-737).
// A batch renders up to GPS_BATCH_MAX_VIEWS independent frames of uneven
// size / spp with one launch per stage for all views: block cull (grid.y =
// view), preprocess (each Gaussian fetched once, evaluated for every view),
// one scan + distribution over the concatenated per-view tiles, one
// persistent splat over the points of all views, and the per-view combine.
// Every view is bit-identical to the same frame rendered alone by a pass
// (occlusion culling off).  The launch sequence can be captured once into a
// HIP graph and replayed for new cameras of the same view layout; the splat
// variant (threads, points per thread) is picked by gps_batch_autotune.
// Views with many samples per point use the binned splat (the regions of all
// such views in one set of launches), the others the point-parallel splat,
// under the same rule as a single pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gps.h"
#include "gps_fb.h"
#include "gps_internal.h"
#include "gps_kernels.h"

using namespace gps;

namespace {

struct Layout {  // what a captured graph depends on
    std::vector<int32_t> shape;  // per view: w, h, ss, spp, hier, regions
    std::vector<const void*> outs;
    void* stream = nullptr;
    int threads = 0, ppt = 0, pre_views = 0;
    bool operator==(const Layout& o) const {
        return shape == o.shape && outs == o.outs && stream == o.stream && threads == o.threads && ppt == o.ppt &&
               pre_views == o.pre_views;
    }
};

}  // namespace

struct gps_batch {
    gps_scene* scene = nullptr;
    gps_batch_settings settings{};
    uint32_t N = 0, npad = 0, tiles_per_view = 0;
    uint64_t fb_view_words = 0;  // FB words reserved per view
    uint64_t bytes = 0;
    BatchView* d_views = nullptr;
    BatchView* h_views = nullptr;  // pinned
    uint32_t* weights = nullptr;
    SplatRecord* recs = nullptr;
    uint32_t *tile_sum = nullptr, *tile_cnt = nullptr, *tile_off = nullptr, *tile_voff = nullptr;
    uint32_t *cidx = nullptr, *coff = nullptr;
    uint8_t* cview = nullptr;
    uint32_t* chunk_start = nullptr;
    PassCounters* ctr = nullptr;   // ctr[0] batch totals; ctr[1].total used as scratch iteration word
    uint32_t* atomics = nullptr;
    CullNode* nodes = nullptr;
    uint32_t* cull_mask = nullptr;
    CullView cull{};
    uint32_t mask_stride = 0;
    bool hier_available = false;
    uint64_t* fb = nullptr;
    uint32_t grid = 0, cus = 256;
    int32_t xcds = 1;
    int threads = 256, ppt = 4;
    int pre_views = 0;  // views per preprocess workgroup (0: all)
    int splat_mode = GPS_SPLAT_AUTO;
    // binned splat buffers, allocated on the first batch with a binned view
    bool bin_ready = false, bin_failed = false;
    BatchBinArgs bin{};
    uint32_t bin_regions_cap = 0;
    bool use_graph = false;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    Layout layout;
    int last_views = 0;
    std::string error;
};

namespace {

template <typename T> gps_status balloc(gps_batch* b, T** p, uint64_t count, bool zero) {
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) return GPS_ERR_OUT_OF_MEMORY;
    b->bytes += count * sizeof(T);
    if (zero && hipMemset(*p, 0, count * sizeof(T)) != hipSuccess) return GPS_ERR_HIP;
    return GPS_OK;
}

void free_binned(gps_batch* b) {
    BatchBinArgs& a = b->bin;
    for (void* p : {(void*)a.erect, (void*)a.ew, (void*)a.big, (void*)a.reg_cnt, (void*)a.reg_pts, (void*)a.reg_ecnt,
                    (void*)a.reg_ptot, (void*)a.reg_eoff, (void*)a.reg_poff, (void*)a.reg_cursor, (void*)a.reg_jobs,
                    (void*)a.job_reg, (void*)a.job_p0, (void*)a.list, (void*)a.lw, (void*)a.lpre, (void*)a.tsum,
                    (void*)a.toff, (void*)a.state})
        if (p) (void)hipFree(p);
    a = BatchBinArgs{};
}

// Binned splat buffers for every compacted entry of a full batch and the
// regions of max_views views of the largest size.  On failure the batch
// keeps rendering every view with the point-parallel splat.
bool ensure_binned(gps_batch* b) {
    if (b->bin_ready) return true;
    if (b->bin_failed) return false;
    const uint64_t V = (uint64_t)b->settings.max_views;
    const uint64_t entries = std::max<uint64_t>(V * b->N, 1);
    uint32_t max_regions = 0;
    for (int spp = 1; spp <= std::min(b->settings.max_spp, 4); ++spp) {
        BinGeom g;
        if (binned_geometry(b->settings.max_width * b->settings.max_supersampling,
                            b->settings.max_height * b->settings.max_supersampling, spp, &g))
            max_regions = std::max(max_regions, g.nregions);
    }
    if (max_regions == 0) {
        b->bin_failed = true;
        return false;
    }
    BatchBinArgs& a = b->bin;
    const uint64_t regions = V * max_regions;
    const uint64_t cap = std::min<uint64_t>(entries + entries / 2 + 65536, 0xFFFFF000ull);
    const uint64_t list_cap = (cap + GPS_TILE - 1) / GPS_TILE * GPS_TILE;
    a.list_cap = (uint32_t)list_cap;
    a.job_cap = (uint32_t)(regions + (1ull << 32) / GPS_BIN_JOB_POINTS + 1);
    a.job_points = GPS_BIN_JOB_POINTS;
    gps_status s = GPS_OK;
    auto chk = [&](gps_status r) {
        if (s == GPS_OK) s = r;
    };
    chk(balloc(b, &a.erect, entries, false));
    chk(balloc(b, &a.ew, entries, false));
    chk(balloc(b, &a.big, entries, false));
    for (uint32_t** p : {&a.reg_cnt, &a.reg_pts, &a.reg_ecnt, &a.reg_ptot, &a.reg_eoff, &a.reg_poff, &a.reg_cursor,
                         &a.reg_jobs})
        chk(balloc(b, p, regions, true));
    chk(balloc(b, &a.job_reg, a.job_cap, false));
    chk(balloc(b, &a.job_p0, a.job_cap, false));
    chk(balloc(b, &a.list, list_cap, false));
    chk(balloc(b, &a.lw, list_cap, true));
    chk(balloc(b, &a.lpre, list_cap, false));
    chk(balloc(b, &a.tsum, list_cap / GPS_TILE + 1, false));
    chk(balloc(b, &a.toff, list_cap / GPS_TILE + 1, false));
    chk(balloc(b, &a.state, 8, true));
    if (s != GPS_OK) {
        free_binned(b);
        (void)hipGetLastError();
        b->bin_failed = true;
        return false;
    }
    b->bin_regions_cap = (uint32_t)regions;
    b->bin_ready = true;
    return true;
}

void drop_graph(gps_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->layout = Layout{};
}

BatchArgs args(gps_batch* b, int nviews) {
    BatchArgs a{};
    a.scene = b->scene->view;
    a.views = b->d_views;
    a.nviews = nviews;
    a.npad = b->npad;
    a.tiles_per_view = b->tiles_per_view;
    a.weights = b->weights;
    a.recs = b->recs;
    a.tile_sum = b->tile_sum;
    a.tile_cnt = b->tile_cnt;
    a.tile_off = b->tile_off;
    a.tile_voff = b->tile_voff;
    a.cidx = b->cidx;
    a.coff = b->coff;
    a.cview = b->cview;
    a.chunk_start = b->chunk_start;
    a.ctr = b->ctr;
    a.iteration = &b->ctr[1].total;
    a.cull = b->cull;
    a.mask_stride = b->mask_stride;
    a.hier = b->hier_available ? 1 : 0;
    a.fb = b->fb;
    a.max_fb_words = b->fb_view_words;
    a.fb_key32 = (uint64_t)(uint32_t)a.nviews * b->fb_view_words < 0xFFFFFFFFull ? 1 : 0;
    a.grid = b->grid;
    a.xcds = b->xcds;
    a.threads = b->threads;
    a.ppt = b->ppt;
    a.pre_views = b->pre_views;
    a.atomics = b->atomics;
    a.bin = b->bin;
    return a;
}

// Every launch of one batch (no host synchronisation, graph-capturable).
gps_status enqueue(gps_batch* b, int nviews, const gps_frame_params* views, float* const* colors,
                   float* const* depths, hipStream_t st) {
    const BatchArgs a = args(b, nviews);
    launch_zero_u32(b->atomics, 1, st);
    launch_batch_clear(a, st);
    if (b->hier_available) launch_batch_cull(a, st);
    launch_batch_preprocess(a, st);
    launch_batch_distribute(a, st);
    launch_batch_binned(a, st);
    if (!launch_batch_splat(a, st)) return GPS_ERR_INVALID_ARG;
    launch_batch_bin_big(a, st);
    for (int v = 0; v < nviews; ++v) {
        const gps_frame_params& f = views[v];
        const int spp = f.samples_per_pixel > 0 ? f.samples_per_pixel : b->settings.max_spp;
        launch_combine(b->fb + (uint64_t)v * b->fb_view_words, colors[v], depths[v], f.width, f.height,
                       f.supersampling_factor, spp, f.near_plane, f.far_plane, 0, st);
    }
    return GPS_OK;
}

}  // namespace

extern "C" {

gps_status gps_batch_create(const gps_batch_settings* st, gps_scene* scene, gps_batch** out) {
    if (!st || !scene || !out) return GPS_ERR_INVALID_ARG;
    *out = nullptr;
    if (st->max_views < 1 || st->max_views > GPS_BATCH_MAX_VIEWS || st->max_width < 1 || st->max_height < 1 ||
        st->max_spp < 1 || st->max_supersampling < 1 || st->max_number_of_points < 0)
        return GPS_ERR_INVALID_ARG;
    if ((int64_t)st->max_width * st->max_supersampling >= 65536 ||
        (int64_t)st->max_height * st->max_supersampling >= 65536)
        return GPS_ERR_UNSUPPORTED;
    gps_batch* b = new gps_batch();
    b->scene = scene;
    b->settings = *st;
    b->N = scene->view.N;
    b->tiles_per_view = (b->N + GPS_TILE - 1) / GPS_TILE;
    b->npad = b->tiles_per_view * GPS_TILE;
    const uint64_t V = (uint64_t)st->max_views;
    if (V * b->npad >= (1ull << 32)) {
        delete b;
        return GPS_ERR_UNSUPPORTED;  // concatenated indices are 32-bit
    }
    b->fb_view_words = fb_words(st->max_width * st->max_supersampling, st->max_height * st->max_supersampling,
                                st->max_spp);
    if (b->fb_view_words >= (1ull << 31)) {
        delete b;
        return GPS_ERR_UNSUPPORTED;  // framebuffer word indices are 32-bit signed
    }
    gps_status s = GPS_OK;
    auto chk = [&](gps_status r) {
        if (s == GPS_OK) s = r;
    };
    chk(balloc(b, &b->d_views, V, true));
    if (s == GPS_OK && hipHostMalloc((void**)&b->h_views, V * sizeof(BatchView), hipHostMallocDefault) != hipSuccess)
        s = GPS_ERR_OUT_OF_MEMORY;
    chk(balloc(b, &b->weights, V * b->npad, true));
    chk(balloc(b, &b->recs, V * b->npad, false));
    chk(balloc(b, &b->tile_sum, V * b->tiles_per_view, true));
    chk(balloc(b, &b->tile_cnt, V * b->tiles_per_view, true));
    chk(balloc(b, &b->tile_off, V * b->tiles_per_view + 1, true));
    chk(balloc(b, &b->tile_voff, V * b->tiles_per_view + 1, true));
    chk(balloc(b, &b->cidx, V * std::max<uint32_t>(b->N, 1), false));
    chk(balloc(b, &b->coff, V * std::max<uint32_t>(b->N, 1), false));
    chk(balloc(b, &b->cview, V * std::max<uint32_t>(b->N, 1), false));
    // The batch's point list is indexed in 32 bits; the smallest splat chunk
    // has 256 points.
    chk(balloc(b, &b->chunk_start, (1ull << 32) / 256 + 2, true));
    chk(balloc(b, &b->ctr, 2, true));
    chk(balloc(b, &b->atomics, 1, true));
    chk(balloc(b, &b->fb, V * b->fb_view_words, false));
    b->hier_available = st->fully_disable_hierarchical_culling == 0 && b->N > 0;
    const uint32_t gpb = st->gaussians_per_block > 0 ? (uint32_t)st->gaussians_per_block
                                                     : (uint32_t)std::max(4.0f, std::floor((float)b->N / (float)1e6));
    const uint32_t nb = (b->N + gpb - 1) / gpb;
    b->cull.gpb = gpb;
    b->cull.node_count = b->hier_available ? nb : 0;
    b->cull.mask_words = (nb + 31) / 32;
    b->mask_stride = b->cull.mask_words + 2;
    if (b->hier_available) {
        chk(balloc(b, &b->nodes, nb, false));
        chk(balloc(b, &b->cull_mask, V * b->mask_stride, true));
        if (s == GPS_OK) launch_build_nodes(scene->view, gpb, nb, b->nodes, nullptr);
    } else {
        chk(balloc(b, &b->cull_mask, V * 2, true));
    }
    b->cull.nodes = b->nodes;
    b->cull.mask = b->cull_mask;
    int cus = 256, dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    b->xcds = (cus % 8 == 0) ? 8 : 1;
    b->grid = (uint32_t)cus * 8u;
    b->cus = (uint32_t)cus;
    if (s == GPS_OK && hipDeviceSynchronize() != hipSuccess) s = GPS_ERR_HIP;
    if (s != GPS_OK) {
        gps_batch_destroy(b);
        return s;
    }
    *out = b;
    return GPS_OK;
}

uint64_t gps_batch_device_bytes(const gps_batch* b) { return b ? b->bytes : 0; }

gps_status gps_batch_render(gps_batch* b, int32_t nviews, const gps_frame_params* views, float* const* colors,
                            float* const* depths, void* stream) {
    if (!b || !views || !colors || !depths || nviews < 1 || nviews > b->settings.max_views) return GPS_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    Layout lay;
    lay.stream = stream;
    lay.threads = b->threads;
    lay.ppt = b->ppt;
    lay.pre_views = b->pre_views;
    uint32_t regions = 0, max_view_regions = 0;
    for (int v = 0; v < nviews; ++v) {
        const gps_frame_params& f = views[v];
        const int spp = f.samples_per_pixel > 0 ? f.samples_per_pixel : b->settings.max_spp;
        if (!colors[v] || !depths[v] || f.width < 1 || f.height < 1 || f.width > b->settings.max_width ||
            f.height > b->settings.max_height || spp < 1 || spp > b->settings.max_spp ||
            f.supersampling_factor < 1 || f.supersampling_factor > b->settings.max_supersampling ||
            !(f.far_plane > f.near_plane))
            return GPS_ERR_INVALID_ARG;
        if (f.enable_occlusion_culling) return GPS_ERR_UNSUPPORTED;  // needs the previous frame of that camera
        BatchView& bv = b->h_views[v];
        std::memset(&bv, 0, sizeof(bv));
        make_frame_consts(&f, spp, &bv.f);
        bv.f.hier = (f.enable_hierarchical_culling && b->hier_available) ? 1 : 0;
        bv.f.occl = 0;
        bv.f.only_prev = 0;
        bv.f.frame_seed = (uint32_t)(4 * f.frame);
        std::memcpy(bv.planes, f.frustum_planes, sizeof(bv.planes));
        bv.fb_offset = (uint64_t)v * b->fb_view_words;
        bv.fb_words = fb_words(f.width * f.supersampling_factor, f.height * f.supersampling_factor, spp);
        bv.background = background_word(f.background[0], f.background[1], f.background[2]);
        bv.budget = (uint32_t)std::min<int64_t>(b->settings.max_number_of_points, 0x7FFFFFFFll);
        // Binned splat for the view under the same rule as a single pass
        // (unbiased frames only: the biased splat adds colours per sample).
        BinGeom g;
        if (bv.f.unbiased && b->splat_mode != GPS_SPLAT_POINT_PARALLEL &&
            (b->splat_mode == GPS_SPLAT_BINNED || bv.f.Ki * bv.f.spp >= GPS_BIN_AUTO_SAMPLES) &&
            binned_geometry(bv.f.up_w, bv.f.up_h, bv.f.spp, &g) && ensure_binned(b) &&
            regions + g.nregions <= b->bin_regions_cap) {
            bv.bg = g;
            bv.rbase = regions;
            bv.dense_points =
                b->splat_mode == GPS_SPLAT_BINNED ? 0xFFFFFFFFu : binned_dense_points(bv.f.up_w, bv.f.up_h, bv.f.Ki);
            regions += g.nregions;
            max_view_regions = std::max(max_view_regions, g.nregions);
        } else {
            bv.rbase = regions;  // bg.nregions == 0: point-parallel
        }
        lay.shape.insert(lay.shape.end(), {f.width, f.height, f.supersampling_factor, spp, bv.f.hier,
                                           (int32_t)bv.bg.nregions});
        lay.outs.push_back(colors[v]);
        lay.outs.push_back(depths[v]);
    }
    b->bin.any = regions > 0 ? 1 : 0;
    b->bin.nregions = regions;
    b->bin.max_view_regions = max_view_regions;
    b->bin.cus = b->cus;
    if (hipMemcpyAsync(b->d_views, b->h_views, nviews * sizeof(BatchView), hipMemcpyHostToDevice, st) != hipSuccess)
        return GPS_ERR_HIP;
    b->last_views = nviews;
    if (!b->use_graph || !stream) return enqueue(b, nviews, views, colors, depths, st);
    if (!(b->exec && lay == b->layout)) {
        drop_graph(b);
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return GPS_ERR_HIP;
        const gps_status e = enqueue(b, nviews, views, colors, depths, st);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(st, &g);
        if (e != GPS_OK || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return e != GPS_OK ? e : GPS_ERR_HIP;
        }
        b->graph = g;
        if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            drop_graph(b);
            return GPS_ERR_HIP;
        }
        b->layout = lay;
    }
    return hipGraphLaunch(b->exec, st) == hipSuccess ? GPS_OK : GPS_ERR_HIP;
}

gps_status gps_batch_set_splat_mode(gps_batch* b, int32_t mode) {
    if (!b || mode < GPS_SPLAT_AUTO || mode > GPS_SPLAT_BINNED) return GPS_ERR_INVALID_ARG;
    if (b->splat_mode != mode) drop_graph(b);
    b->splat_mode = mode;
    return GPS_OK;
}

gps_status gps_batch_set_tuning(gps_batch* b, const gps_batch_tuning* t) {
    if (!b || !t) return GPS_ERR_INVALID_ARG;
    bool ok = false;
    for (int i = 0; i < kNumBatchSplatVariants; ++i)
        ok |= kBatchSplatVariants[i][0] == t->splat_threads && kBatchSplatVariants[i][1] == t->splat_points_per_thread;
    if (!ok || t->preprocess_views_per_workgroup < 0 || t->preprocess_views_per_workgroup > GPS_BATCH_MAX_VIEWS)
        return GPS_ERR_INVALID_ARG;
    if (b->threads != t->splat_threads || b->ppt != t->splat_points_per_thread ||
        b->pre_views != t->preprocess_views_per_workgroup)
        drop_graph(b);
    b->threads = t->splat_threads;
    b->ppt = t->splat_points_per_thread;
    b->pre_views = t->preprocess_views_per_workgroup;
    b->use_graph = t->use_graph != 0;
    return GPS_OK;
}

gps_status gps_batch_autotune(gps_batch* b, int32_t nviews, const gps_frame_params* views, float* const* colors,
                              float* const* depths, void* stream, int32_t repeats, gps_batch_tuning* out) {
    if (!b || !out || repeats < 1) return GPS_ERR_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    std::memset(out, 0, sizeof(*out));
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return GPS_ERR_HIP;
    const bool graph = b->use_graph;
    b->use_graph = false;
    double best = 1e30;
    int best_i = 0;
    gps_status s = GPS_OK;
    auto time_batch = [&](float* ms) {
        gps_status r = gps_batch_render(b, nviews, views, colors, depths, stream);  // warm-up
        *ms = 0.0f;
        if (r == GPS_OK && hipEventRecord(e0, st) != hipSuccess) r = GPS_ERR_HIP;
        for (int k = 0; k < repeats && r == GPS_OK; ++k) r = gps_batch_render(b, nviews, views, colors, depths, stream);
        if (r == GPS_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(ms, e0, e1) != hipSuccess))
            r = GPS_ERR_HIP;
        return r;
    };
    // splat variants at the current preprocess split, then the preprocess
    // splits with the fastest splat variant
    for (int i = 0; i < kNumBatchSplatVariants && s == GPS_OK; ++i) {
        b->threads = kBatchSplatVariants[i][0];
        b->ppt = kBatchSplatVariants[i][1];
        float ms = 0.0f;
        s = time_batch(&ms);
        out->variant_ms[i] = ms / repeats;
        out->variant_threads[i] = kBatchSplatVariants[i][0];
        out->variant_points_per_thread[i] = kBatchSplatVariants[i][1];
        if (s == GPS_OK && ms / repeats < best) {
            best = ms / repeats;
            best_i = i;
        }
    }
    out->variants = kNumBatchSplatVariants;
    out->splat_threads = kBatchSplatVariants[best_i][0];
    out->splat_points_per_thread = kBatchSplatVariants[best_i][1];
    b->threads = out->splat_threads;
    b->ppt = out->splat_points_per_thread;
    int best_v = b->pre_views;
    for (int i = 0; i < kNumBatchPreViews && s == GPS_OK; ++i) {
        b->pre_views = kBatchPreViews[i];
        float ms = 0.0f;
        s = time_batch(&ms);
        out->pre_variant_views[i] = kBatchPreViews[i];
        out->pre_variant_ms[i] = ms / repeats;
        if (s == GPS_OK && ms / repeats < best) {
            best = ms / repeats;
            best_v = kBatchPreViews[i];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out->pre_variants = kNumBatchPreViews;
    out->preprocess_views_per_workgroup = best_v;
    out->use_graph = graph ? 1 : 0;
    b->pre_views = best_v;
    b->use_graph = graph;
    drop_graph(b);
    return s;
}

gps_status gps_batch_view_stats(gps_batch* b, int32_t view, gps_stats* out, void* stream) {
    if (!b || !out || view < 0 || view >= b->last_views) return GPS_ERR_INVALID_ARG;
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return GPS_ERR_HIP;
    BatchView bv;
    if (hipMemcpy(&bv, b->d_views + view, sizeof(bv), hipMemcpyDeviceToHost) != hipSuccess) return GPS_ERR_HIP;
    std::memset(out, 0, sizeof(*out));
    out->num_gaussians = b->N;
    out->passes = 1;
    out->K = bv.f.Ki;
    out->valid_gaussians[0] = bv.valid;
    out->total_weight[0] = bv.total;
    out->num_points[0] = bv.points;
    out->culled_blocks[0] = bv.culled_blocks;
    out->samples = (uint64_t)bv.points * (uint64_t)bv.f.Ki * (uint64_t)bv.f.spp;
    out->budget_truncated = bv.total > bv.points ? 1 : 0;
    if (b->hier_available) {
        out->gaussians_per_block = (int32_t)b->cull.gpb;
        out->num_blocks = b->cull.node_count;
    }
    uint32_t atomics = 0;
    if (hipMemcpy(&atomics, b->atomics, 4, hipMemcpyDeviceToHost) != hipSuccess) return GPS_ERR_HIP;
    out->atomics = atomics;  // whole batch
    if (bv.bg.nregions) {
        uint32_t st[8];
        if (hipMemcpy(st, b->bin.state, sizeof(st), hipMemcpyDeviceToHost) != hipSuccess) return GPS_ERR_HIP;
        const bool dense = bv.points > bv.dense_points;
        out->splat_binned = st[3] || dense ? 0 : 1;
        out->splat_fallback = st[3] ? 1 : dense ? 2 : 0;
        out->bin_entries = st[0];  // whole batch
        out->bin_jobs = st[1];
        out->bin_big = st[2];
    }
    return GPS_OK;
}

void gps_batch_destroy(gps_batch* b) {
    if (!b) return;
    drop_graph(b);
    (void)hipFree(b->d_views);
    if (b->h_views) (void)hipHostFree(b->h_views);
    (void)hipFree(b->weights);
    (void)hipFree(b->recs);
    (void)hipFree(b->tile_sum);
    (void)hipFree(b->tile_cnt);
    (void)hipFree(b->tile_off);
    (void)hipFree(b->tile_voff);
    (void)hipFree(b->cidx);
    (void)hipFree(b->coff);
    (void)hipFree(b->cview);
    (void)hipFree(b->chunk_start);
    (void)hipFree(b->ctr);
    (void)hipFree(b->atomics);
    (void)hipFree(b->nodes);
    (void)hipFree(b->cull_mask);
    (void)hipFree(b->fb);
    free_binned(b);
    delete b;
}

}  // extern "C"
