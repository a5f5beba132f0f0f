// gps_api.cpp — host orchestration and the C ABI of include/gps.h.
//
// A gps_pass owns every per-frame device buffer; gps_pass_render() only
// enqueues kernels on the caller's stream (no host synchronization, no
// allocation unless the supersampling factor changed), replacing
// GaussianPointSplatting::render (gaussian_point_splatting.cu:599-737) whose
// workload distributor read three scalars back to the host per pass
// (workload_distributor.cu:70-107).  Here the point count stays on the device
// and the splat kernel is persistent over a device-side chunk count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gps.h"
#include "gps_fb.h"
#include "gps_hostmath.h"
#include "gps_internal.h"
#include "gps_kernels.h"

using namespace gps;

namespace {

thread_local std::string g_last_error;

gps_status hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? GPS_ERR_OUT_OF_MEMORY : GPS_ERR_HIP;
}

#define GPS_HIP(call)                                       \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(e_, #call);   \
    } while (0)

template <typename T> gps_status dalloc(T** p, uint64_t count, bool zero) {
    *p = nullptr;
    if (count == 0) count = 1;
    GPS_HIP(hipMalloc((void**)p, count * sizeof(T)));
    if (zero) GPS_HIP(hipMemset(*p, 0, count * sizeof(T)));
    return GPS_OK;
}
template <typename T> void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }

}  // namespace

uint16_t gps_float_to_half_rne(float f);
void gps_encode_sh_record(const float* sh, int ncoef, uint8_t out[64]);
void gps_encode_quantized_one(const gps_scene_desc* d, const gps_quantized_bounds* b, uint64_t g, uint32_t geo[4],
                              uint8_t* opacity, uint8_t rec[64]);

namespace {


}  // namespace

struct gps_pass {
    gps_scene* scene = nullptr;
    gps_render_settings settings{};
    uint32_t N = 0;
    uint32_t budget = 0;
    // distributor
    uint32_t* weights = nullptr;
    uint32_t* maskA = nullptr;
    uint32_t* maskB = nullptr;
    SplatRecord* recs = nullptr;
    uint32_t ntiles = 0;
    uint32_t *tile_sum = nullptr, *tile_cnt = nullptr, *tile_off = nullptr, *tile_voff = nullptr;
    uint32_t *cidx = nullptr, *coff = nullptr, *chunk_start = nullptr;
    FrameCounters* ctr = nullptr;
    uint32_t* dbg_indices = nullptr;
    // binned splat
    int32_t splat_mode = GPS_SPLAT_AUTO;
    int32_t count_atomics = 0;
    int32_t count_rejects = 0;
    uint32_t cus = 256;
    uint2* erect = nullptr;
    uint32_t* ew = nullptr;
    uint32_t *reg_cnt = nullptr, *reg_pts = nullptr, *reg_ecnt = nullptr, *reg_ptot = nullptr, *reg_eoff = nullptr,
             *reg_poff = nullptr, *reg_cursor = nullptr, *reg_jobs = nullptr;
    uint32_t *job_reg = nullptr, *job_p0 = nullptr;
    uint32_t *list = nullptr, *lw = nullptr, *lpre = nullptr;
    uint32_t *tsum = nullptr, *toff = nullptr;
    uint32_t* big = nullptr;
    uint32_t* bin_state = nullptr;
    uint32_t list_cap = 0, job_cap = 0;
    int last_binned = 0;
    uint64_t dbg_indices_cap = 0;
    // framebuffers
    uint64_t* fbA = nullptr;
    uint64_t* fbB = nullptr;
    uint64_t fb_words = 0;  // words of the current frame
    uint64_t fb_cap = 0;    // words allocated per buffer
    int fb_ss = -1;
    uint64_t* cur_fb = nullptr;
    // occlusion
    HiZView hiz{};
    uint32_t hiz_size = 0;
    CullNode* nodes = nullptr;
    uint32_t* cull_mask = nullptr;
    CullView cull{};
    bool hier_available = false;
    GroupBox* gboxes = nullptr;  // group occlusion test (null: off)
    uint8_t* gocc = nullptr;     // occluded-group bytes (padded to whole words)
    uint32_t* live = nullptr;    // unculled blocks of the pass, listed by the block cull
    uint8_t* wg_live = nullptr;  // per preprocess workgroup: its block is unculled (this pass)
    GroupView group{};
    int32_t group_culling = 1;
    // splat grid
    uint32_t splat_grid = 0;
    int32_t xcds = 1;
    // ordered point-parallel splat (dense passes)
    int32_t ordered_mode = GPS_ORDERED_AUTO;
    uint16_t* ord_bucket = nullptr;  // per chunk
    uint32_t* ord_hist = nullptr;    // GPS_ORD_BINS (zero between passes)
    uint32_t* ord_lim = nullptr;     // GPS_ORD_PHASES
    uint32_t* ord_tmax = nullptr;    // per 16x16 tile of the framebuffer allocation
    // stage timing: ring of frames x stages x 2 occurrences x (start, stop)
    std::vector<hipEvent_t> ev;
    std::vector<uint8_t> ev_used;
    int timing_cap = 0;
    uint64_t timing_frames = 0;
    int timing_slot = -1;
    // last frame
    int last_occl = 0;
    int last_hier = 0;
    int last_K = 1;
    int last_spp = 1;
};

namespace {

bool camera_finite(const FrameConsts& f) {
    for (int i = 0; i < 16; ++i)
        if (!std::isfinite(f.V[i]) || !std::isfinite(f.VP[i])) return false;
    return std::isfinite(f.focal[0]) && std::isfinite(f.focal[1]) && std::isfinite(f.tan_fov[0]) &&
           std::isfinite(f.tan_fov[1]) && std::isfinite(f.dilation);
}

uint64_t mip_total(HiZView& h, int w, int hgt) {
    h.width = w;
    h.height = hgt;
    int maxd = std::max(w, hgt), e = 0;
    while ((1 << e) < maxd) ++e;
    h.levels = e + 1;
    uint64_t total = 0;
    for (int l = 0; l < h.levels; ++l) {
        const int fe = h.levels - 1;
        h.lw[l] = std::max(1, (w + (1 << (fe - l)) - 1) >> (fe - l));
        h.lh[l] = std::max(1, (hgt + (1 << (fe - l)) - 1) >> (fe - l));
        h.off[l] = (uint32_t)total;
        total += (uint64_t)h.lw[l] * h.lh[l];
    }
    return total;
}

// Allocates FB A/B for frames of supersampling `ss` up to the pass's size and
// spp (reallocated only when ss changes); `frame_words` is this frame's part.
gps_status ensure_fb(gps_pass* p, int ss, uint64_t frame_words) {
    const int uw = p->settings.width * ss, uh = p->settings.height * ss;
    const uint64_t words = fb_words(uw, uh, p->settings.render_pass_count);
    if (words == p->fb_cap && p->fbA) {
        p->fb_words = frame_words;
        return GPS_OK;
    }
    if (words >= (1ull << 31)) {
        g_last_error = "framebuffer exceeds 2^31 words";
        return GPS_ERR_UNSUPPORTED;
    }
    dfree(p->fbA);
    dfree(p->fbB);
    dfree(p->ord_tmax);
    gps_status s = dalloc(&p->fbA, words, false);
    if (s != GPS_OK) return s;
    s = dalloc(&p->fbB, words, false);
    if (s != GPS_OK) return s;
    // one word per 16x16 tile of a 1-layer framebuffer of the allocation's size
    s = dalloc(&p->ord_tmax, words / (256ull * (uint64_t)p->settings.render_pass_count), false);
    if (s != GPS_OK) return s;
    p->fb_cap = words;
    p->fb_words = frame_words;
    p->fb_ss = ss;
    return GPS_OK;
}

constexpr int kOcc = 2;
inline int ev_index(int slot, int stage, int occ, int which) {
    return ((slot * GPS_NUM_STAGES + stage) * kOcc + occ) * 2 + which;
}

// Records the start (which = 0) or stop (1) event of `stage`, occurrence `occ`
// of the current frame's slot.
void stage_mark(gps_pass* p, int stage, int occ, int which, hipStream_t st) {
    if (p->timing_slot < 0) return;
    const int i = ev_index(p->timing_slot, stage, occ, which);
    (void)hipEventRecord(p->ev[i], st);
    if (which == 1) p->ev_used[i / 2] = 1;
}

void destroy_events(gps_pass* p) {
    for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
    p->ev.clear();
    p->ev_used.clear();
    p->timing_cap = 0;
    p->timing_slot = -1;
    p->timing_frames = 0;
}

}  // namespace

extern "C" {

int32_t gps_abi_version(void) { return GPS_ABI_VERSION; }

const char* gps_status_str(gps_status s) {
    switch (s) {
        case GPS_OK: return "ok";
        case GPS_ERR_INVALID_ARG: return "invalid argument";
        case GPS_ERR_OUT_OF_MEMORY: return "out of device memory";
        case GPS_ERR_HIP: return "HIP runtime error";
        case GPS_ERR_UNSUPPORTED: return "unsupported";
        case GPS_ERR_IO: return "I/O error";
        case GPS_ERR_NOT_READY: return "not ready";
    }
    return "unknown status";
}

const char* gps_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------ scenes
extern "C++" {
namespace {

// Runs fn(lo, hi) over [0, n) on up to 32 host threads.
template <typename F> void parallel_range(uint64_t n, F fn) {
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt == 0 ? 1 : std::min(nt, 32u);
    if (n < 16384) nt = 1;
    std::vector<std::thread> ts;
    const uint64_t per = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const uint64_t lo = (uint64_t)t * per;
        if (lo >= n) break;
        ts.emplace_back(fn, lo, std::min(n, lo + per));
    }
    for (auto& t : ts) t.join();
}

}  // namespace
}

gps_status gps_scene_create_empty(uint64_t N, int32_t sh_degree, int32_t format, gps_scene** out) {
    if (!out) return GPS_ERR_INVALID_ARG;
    *out = nullptr;
    if (sh_degree < 0 || sh_degree > 3) return GPS_ERR_INVALID_ARG;
    if (format != GPS_FORMAT_FP32 && format != GPS_FORMAT_HALF_FP8 && format != GPS_FORMAT_QUANTIZED)
        return GPS_ERR_INVALID_ARG;
    if (N >= (1ull << 32) - GPS_TILE) return GPS_ERR_UNSUPPORTED;
    const int ncoef = (sh_degree + 1) * (sh_degree + 1);
    const uint64_t stride = round_up(std::max<uint64_t>(N, 1), 64);
    gps_scene* s = new gps_scene();
    s->view.N = (uint32_t)N;
    s->view.sh_degree = sh_degree;
    s->view.format = format;
    s->view.stride = stride;
    gps_status st;
    if (format == GPS_FORMAT_FP32) {
        const int nplanes = 11 + 3 * ncoef;
        float* planes = nullptr;
        st = dalloc(&planes, stride * nplanes, true);
        s->planes = planes;
        s->view.planes = planes;
        s->bytes = stride * nplanes * sizeof(float);
    } else if (format == GPS_FORMAT_QUANTIZED) {
        uint4* geo = nullptr;
        uint4* shrec = nullptr;
        uint8_t* qop = nullptr;
        float* qb = nullptr;
        st = dalloc(&geo, stride, true);
        if (st == GPS_OK) st = dalloc(&shrec, 4 * stride, true);
        if (st == GPS_OK) st = dalloc(&qop, stride, true);
        if (st == GPS_OK) st = dalloc(&qb, sizeof(gps_quantized_bounds) / sizeof(float), true);
        s->geo = geo;
        s->shrec = shrec;
        s->qop = qop;
        s->qbounds = qb;
        s->view.geo = geo;
        s->view.shrec = shrec;
        s->view.qop = qop;
        s->view.qbounds = qb;
        s->bytes = stride * (16 + 64 + 1) + sizeof(gps_quantized_bounds);
    } else {
        float* means = nullptr;
        uint4* geo = nullptr;
        uint4* shrec = nullptr;
        st = dalloc(&means, 3 * stride, true);
        if (st == GPS_OK) st = dalloc(&geo, stride, true);
        if (st == GPS_OK) st = dalloc(&shrec, 4 * stride, true);
        s->planes = means;
        s->geo = geo;
        s->shrec = shrec;
        s->view.means = means;
        s->view.geo = geo;
        s->view.shrec = shrec;
        s->bytes = stride * (12 + 16 + 64);
    }
    if (st != GPS_OK) {
        gps_scene_destroy(s);
        return st;
    }
    *out = s;
    return GPS_OK;
}

gps_status gps_scene_upload_range(gps_scene* s, uint64_t first, const gps_scene_desc* d) {
    if (!s || !d) return GPS_ERR_INVALID_ARG;
    const uint64_t n_all = d->num_gaussians;
    if (d->sh_degree != s->view.sh_degree || d->format != s->view.format) return GPS_ERR_INVALID_ARG;
    if (first > s->view.N || n_all > s->view.N - first) return GPS_ERR_INVALID_ARG;
    if (n_all > 0 && (!d->means || !d->rotations || !d->scales || !d->opacities || !d->shs))
        return GPS_ERR_INVALID_ARG;
    const int ncoef = (d->sh_degree + 1) * (d->sh_degree + 1);
    const uint64_t stride = s->view.stride;
    const uint64_t chunk = 1u << 20;
    if (d->format == GPS_FORMAT_FP32) {
        const int nplanes = 11 + 3 * ncoef;
        float* planes = (float*)s->planes;
        std::vector<float> stage(chunk * (uint64_t)nplanes);
        for (uint64_t b = 0; b < n_all; b += chunk) {
            const uint64_t n = std::min(chunk, n_all - b);
            parallel_range(n, [&](uint64_t lo, uint64_t hi) {
                for (uint64_t i = lo; i < hi; ++i) {
                    const uint64_t g = b + i;
                    for (int c = 0; c < 3; ++c) stage[c * chunk + i] = d->means[3 * g + c];
                    for (int c = 0; c < 4; ++c) stage[(3 + c) * chunk + i] = d->rotations[4 * g + c];
                    for (int c = 0; c < 3; ++c) stage[(7 + c) * chunk + i] = d->scales[3 * g + c];
                    stage[10 * chunk + i] = d->opacities[g];
                    for (int c = 0; c < 3 * ncoef; ++c) stage[(11 + c) * chunk + i] = d->shs[(uint64_t)3 * ncoef * g + c];
                }
            });
            for (int p = 0; p < nplanes; ++p) {
                hipError_t e = hipMemcpy(planes + (uint64_t)p * stride + first + b, stage.data() + p * chunk,
                                         n * sizeof(float), hipMemcpyHostToDevice);
                if (e != hipSuccess) return hip_fail(e, "scene upload");
            }
        }
    } else if (d->format == GPS_FORMAT_QUANTIZED) {
        // The reference's compressed scene, encoded on the host against the
        // scene box set beforehand.
        if (!s->have_bounds) return GPS_ERR_INVALID_ARG;
        uint4* geo = (uint4*)s->geo;
        uint4* shrec = (uint4*)s->shrec;
        uint8_t* qop = (uint8_t*)s->qop;
        std::vector<uint32_t> gstage(4 * chunk), sstage(16 * chunk);
        std::vector<uint8_t> ostage(chunk);
        for (uint64_t b = 0; b < n_all; b += chunk) {
            const uint64_t n = std::min(chunk, n_all - b);
            parallel_range(n, [&](uint64_t lo, uint64_t hi) {
                for (uint64_t i = lo; i < hi; ++i)
                    gps_encode_quantized_one(d, &s->bounds, b + i, &gstage[4 * i], &ostage[i],
                                             reinterpret_cast<uint8_t*>(&sstage[16 * i]));
            });
            hipError_t e = hipMemcpy(geo + first + b, gstage.data(), n * 16, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(shrec + 4 * (first + b), sstage.data(), n * 64, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(qop + first + b, ostage.data(), n, hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(e, "scene upload");
        }
    } else {
        // GPS_FORMAT_HALF_FP8: encoded on the host (fp16 RNE, fp8 E4M3 RNE with a
        // per-Gaussian scale = max|rest coefficient| / 448).
        float* means = (float*)s->planes;
        uint4* geo = (uint4*)s->geo;
        uint4* shrec = (uint4*)s->shrec;
        std::vector<float> mstage(3 * chunk);
        std::vector<uint32_t> gstage(4 * chunk), sstage(16 * chunk);
        for (uint64_t b = 0; b < n_all; b += chunk) {
            const uint64_t n = std::min(chunk, n_all - b);
            parallel_range(n, [&](uint64_t lo, uint64_t hi) {
                uint8_t enc[64];
                for (uint64_t i = lo; i < hi; ++i) {
                    const uint64_t g = b + i;
                    for (int c = 0; c < 3; ++c) mstage[c * chunk + i] = d->means[3 * g + c];
                    uint16_t h[8];
                    for (int k = 0; k < 4; ++k) h[k] = gps_float_to_half_rne(d->rotations[4 * g + k]);
                    for (int k = 0; k < 3; ++k) h[4 + k] = gps_float_to_half_rne(d->scales[3 * g + k]);
                    h[7] = gps_float_to_half_rne(d->opacities[g]);
                    std::memcpy(&gstage[4 * i], h, 16);
                    gps_encode_sh_record(d->shs + (uint64_t)3 * ncoef * g, ncoef, enc);
                    std::memcpy(&sstage[16 * i], enc, 64);
                }
            });
            hipError_t e = hipSuccess;
            for (int c = 0; c < 3 && e == hipSuccess; ++c)
                e = hipMemcpy(means + (uint64_t)c * stride + first + b, mstage.data() + c * chunk, n * 4,
                              hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(geo + first + b, gstage.data(), n * 16, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(shrec + 4 * (first + b), sstage.data(), n * 64, hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(e, "scene upload");
        }
    }
    return GPS_OK;
}

gps_status gps_scene_create(const gps_scene_desc* d, gps_scene** out) {
    if (!d || !out) return GPS_ERR_INVALID_ARG;
    *out = nullptr;
    if (d->sh_degree < 0 || d->sh_degree > 3) return GPS_ERR_INVALID_ARG;
    if (d->num_gaussians > 0 && (!d->means || !d->rotations || !d->scales || !d->opacities || !d->shs))
        return GPS_ERR_INVALID_ARG;
    gps_scene* s = nullptr;
    gps_status st = gps_scene_create_empty(d->num_gaussians, d->sh_degree, d->format, &s);
    if (st != GPS_OK) return st;
    if (d->format == GPS_FORMAT_QUANTIZED) {
        gps_quantized_bounds b;
        st = gps_quantized_bounds_compute(d, &b);
        if (st == GPS_OK) st = gps_scene_set_quantized_bounds(s, &b);
    }
    if (st == GPS_OK) st = gps_scene_upload_range(s, 0, d);
    if (st != GPS_OK) {
        gps_scene_destroy(s);
        return st;
    }
    *out = s;
    return GPS_OK;
}

void gps_scene_destroy(gps_scene* s) {
    if (!s) return;
    if (s->planes) (void)hipFree(s->planes);
    if (s->geo) (void)hipFree(s->geo);
    if (s->shrec) (void)hipFree(s->shrec);
    if (s->qop) (void)hipFree(s->qop);
    if (s->qbounds) (void)hipFree(s->qbounds);
    delete s;
}

gps_status gps_scene_set_quantized_bounds(gps_scene* s, const gps_quantized_bounds* b) {
    if (!s || !b) return GPS_ERR_INVALID_ARG;
    if (s->view.format != GPS_FORMAT_QUANTIZED) return GPS_ERR_INVALID_ARG;
    s->bounds = *b;
    s->have_bounds = true;
    const hipError_t e = hipMemcpy(s->qbounds, b, sizeof(*b), hipMemcpyHostToDevice);
    return e == hipSuccess ? GPS_OK : hip_fail(e, "scene bounds");
}

uint64_t gps_scene_device_bytes(const gps_scene* s) { return s ? s->bytes : 0; }

// ------------------------------------------------------------------ passes
gps_status gps_pass_create(const gps_render_settings* st, gps_scene* scene, gps_pass** out) {
    if (!st || !scene || !out) return GPS_ERR_INVALID_ARG;
    *out = nullptr;
    if (st->width <= 0 || st->height <= 0 || st->render_pass_count < 1 || st->supersampling_factor < 1 ||
        st->max_number_of_points < 0)
        return GPS_ERR_INVALID_ARG;
    if ((int64_t)st->width * st->supersampling_factor >= 65536 || (int64_t)st->height * st->supersampling_factor >= 65536)
        return GPS_ERR_UNSUPPORTED;
    // framebuffer word indices are 32-bit signed (fb_index)
    if (fb_words(st->width * st->supersampling_factor, st->height * st->supersampling_factor,
                 st->render_pass_count) >= (1ull << 31))
        return GPS_ERR_UNSUPPORTED;
    gps_pass* p = new gps_pass();
    p->scene = scene;
    p->settings = *st;
    p->N = scene->view.N;
    p->budget = (uint32_t)std::min<int64_t>(st->max_number_of_points, 0x7FFFFFFFll);
    const uint64_t N = p->N;
    p->ntiles = (uint32_t)((N + GPS_TILE - 1) / GPS_TILE);
    const uint64_t npad = (uint64_t)p->ntiles * GPS_TILE;
    const uint64_t nwords = (N + 31) / 32 + 2;
    gps_status s = GPS_OK;
    auto chk = [&](gps_status r) {
        if (s == GPS_OK) s = r;
    };
    chk(dalloc(&p->weights, npad, true));
    chk(dalloc(&p->maskA, nwords, true));
    chk(dalloc(&p->maskB, nwords, true));
    chk(dalloc(&p->recs, N, false));
    chk(dalloc(&p->tile_sum, p->ntiles, true));
    chk(dalloc(&p->tile_cnt, p->ntiles, true));
    chk(dalloc(&p->tile_off, p->ntiles, true));
    chk(dalloc(&p->tile_voff, p->ntiles, true));
    chk(dalloc(&p->cidx, N, false));
    chk(dalloc(&p->coff, N, false));
    chk(dalloc(&p->chunk_start, (uint64_t)p->budget / GPS_SPLAT_CHUNK + 2, true));
    chk(dalloc(&p->ord_bucket, (uint64_t)p->budget / GPS_SPLAT_CHUNK + 2, false));
    chk(dalloc(&p->ord_hist, GPS_ORD_BINS, true));
    chk(dalloc(&p->ord_lim, GPS_ORD_PHASES, true));
    // Binned splat lists: up to 2 regions per valid Gaussian on average
    // (overflow falls back to the point-parallel splat on the device).
    {
        const uint64_t cap = round_up(std::min<uint64_t>(2 * std::max<uint64_t>(N, 1) + 65536, 0xFFFFF000ull), GPS_TILE);
        p->list_cap = (uint32_t)cap;
        p->job_cap = GPS_BIN_HIST + p->budget / GPS_BIN_JOB_POINTS + 1;
        chk(dalloc(&p->erect, std::max<uint64_t>(N, 1), false));
        chk(dalloc(&p->ew, std::max<uint64_t>(N, 1), false));
        chk(dalloc(&p->big, std::max<uint64_t>(N, 1), false));
        for (uint32_t** a : {&p->reg_cnt, &p->reg_pts, &p->reg_ecnt, &p->reg_ptot, &p->reg_eoff, &p->reg_poff,
                             &p->reg_cursor, &p->reg_jobs})
            chk(dalloc(a, GPS_BIN_HIST, true));
        chk(dalloc(&p->job_reg, p->job_cap, false));
        chk(dalloc(&p->job_p0, p->job_cap, false));
        chk(dalloc(&p->list, cap, false));
        chk(dalloc(&p->lw, cap, true));
        chk(dalloc(&p->lpre, cap, false));
        chk(dalloc(&p->tsum, cap / GPS_TILE + 1, false));
        chk(dalloc(&p->toff, cap / GPS_TILE + 1, false));
        chk(dalloc(&p->bin_state, 8, true));
    }
    chk(dalloc(&p->ctr, 1, true));
    // Hi-Z pyramid at output resolution, initialised to +inf (nothing occluded).
    p->hiz_size = (uint32_t)mip_total(p->hiz, st->width, st->height);
    chk(dalloc(&p->hiz.mip, p->hiz_size, false));
    if (s == GPS_OK) launch_fill_f32(p->hiz.mip, p->hiz_size, __builtin_inff(), nullptr);
    // Culling blocks (gaussian_bvh.cu:171-195).
    p->hier_available = st->fully_disable_hierarchical_culling == 0;
    uint32_t gpb = st->gaussians_per_block > 0 ? (uint32_t)st->gaussians_per_block
                                               : (uint32_t)std::max(4.0f, std::floor((float)N / (float)1e6));
    const uint32_t nb = (uint32_t)((N + gpb - 1) / gpb);
    p->cull.gpb = gpb;
    p->cull.node_count = nb;
    p->cull.mask_words = (nb + 31) / 32;
    if (p->hier_available) {
        chk(dalloc(&p->nodes, nb, false));
        chk(dalloc(&p->cull_mask, p->cull.mask_words + 2, true));
        if (s == GPS_OK) launch_build_nodes(scene->view, gpb, nb, p->nodes, nullptr);
    } else {
        chk(dalloc(&p->cull_mask, 2, true));
    }
    p->cull.nodes = p->nodes;
    p->cull.mask = p->cull_mask;
    // Group bounds for the group occlusion test (0.75 byte per Gaussian).
    p->group.count = (uint32_t)((N + GPS_GROUP - 1) / GPS_GROUP);
    if (N > 0) {
        chk(dalloc(&p->gboxes, p->group.count, false));
        chk(dalloc(&p->gocc, round_up(p->group.count, 4) + 4, true));
        if (p->hier_available) {
            chk(dalloc(&p->live, nb, false));
            chk(dalloc(&p->wg_live, (N + GPS_PRE_THREADS - 1) / GPS_PRE_THREADS, false));
        }
        if (s == GPS_OK) launch_group_boxes(scene->view, p->gboxes, p->group.count, nullptr);
    }
    p->group.boxes = p->gboxes;
    p->group.occluded = p->gocc;
    if (s == GPS_OK)
        s = ensure_fb(p, st->supersampling_factor,
                      fb_words(st->width * st->supersampling_factor, st->height * st->supersampling_factor,
                               st->render_pass_count));
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    // MI300-class parts (gfx942 / gfx950) have 8 XCDs of 32 or 38 CUs; the
    // chunk schedule only needs the count to be a divisor of the grid.
    p->xcds = (cus % 8 == 0) ? 8 : 1;
    p->splat_grid = (uint32_t)cus * GPS_SPLAT_WGS_PER_CU;  // persistent; finer chunk interleave than one wave per SIMD slot
    p->cus = (uint32_t)cus;
    if (s == GPS_OK) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) s = hip_fail(e, "pass setup");
    }
    if (s != GPS_OK) {
        gps_pass_destroy(p);
        return s;
    }
    *out = p;
    return GPS_OK;
}

gps_status gps_pass_set_ordered_splat(gps_pass* p, int32_t mode) {
    if (!p || (mode != GPS_ORDERED_AUTO && mode != GPS_ORDERED_OFF && mode != GPS_ORDERED_ALWAYS))
        return GPS_ERR_INVALID_ARG;
    p->ordered_mode = mode;
    return GPS_OK;
}

gps_status gps_pass_set_splat_mode(gps_pass* p, int32_t mode) {
    if (!p || (mode != GPS_SPLAT_AUTO && mode != GPS_SPLAT_POINT_PARALLEL && mode != GPS_SPLAT_BINNED))
        return GPS_ERR_INVALID_ARG;
    p->splat_mode = mode;
    return GPS_OK;
}

gps_status gps_pass_set_group_culling(gps_pass* p, int32_t enable) {
    if (!p || (enable != 0 && enable != 1)) return GPS_ERR_INVALID_ARG;
    p->group_culling = enable;
    return GPS_OK;
}

gps_status gps_pass_set_counters(gps_pass* p, int32_t flags) {
    if (!p || (flags & ~(GPS_COUNT_ATOMICS | GPS_COUNT_REJECTS))) return GPS_ERR_INVALID_ARG;
    p->count_atomics = (flags & GPS_COUNT_ATOMICS) ? 1 : 0;
    p->count_rejects = (flags & GPS_COUNT_REJECTS) ? 1 : 0;
    return GPS_OK;
}

uint64_t gps_pass_estimate_bytes(const gps_render_settings* st, const gps_scene* scene) {
    if (!st || !scene) return 0;
    const uint64_t N = scene->view.N;
    const uint64_t ntiles = (N + GPS_TILE - 1) / GPS_TILE, npad = ntiles * GPS_TILE, nwords = (N + 31) / 32 + 2;
    const uint64_t budget = (uint64_t)std::min<int64_t>(st->max_number_of_points, 0x7FFFFFFFll);
    HiZView h{};
    const uint64_t hiz = mip_total(h, st->width, st->height);
    const uint64_t fb = fb_words(st->width * st->supersampling_factor, st->height * st->supersampling_factor,
                                 st->render_pass_count);
    uint64_t b = npad * 4 + 2 * nwords * 4 + std::max<uint64_t>(N, 1) * (sizeof(SplatRecord) + 8) +
                 16 * ntiles + (budget / GPS_SPLAT_CHUNK + 2) * 4 + sizeof(FrameCounters) + hiz * 4 + 2 * fb * 8;
    b += (budget / GPS_SPLAT_CHUNK + 2) * 2 + 4ull * (GPS_ORD_BINS + GPS_ORD_PHASES) +
         fb / (256ull * (uint64_t)st->render_pass_count) * 4;  // ordered splat
    {
        const uint64_t cap = round_up(std::min<uint64_t>(2 * std::max<uint64_t>(N, 1) + 65536, 0xFFFFF000ull), GPS_TILE);
        b += std::max<uint64_t>(N, 1) * 16 + 3 * cap * 4 + 2 * (cap / GPS_TILE + 1) * 4 + 8ull * 4 * GPS_BIN_HIST +
             2ull * 4 * (GPS_BIN_HIST + budget / GPS_BIN_JOB_POINTS + 1) + 32;
    }
    if (!st->fully_disable_hierarchical_culling) {
        const uint64_t gpb = st->gaussians_per_block > 0 ? (uint64_t)st->gaussians_per_block
                                                         : (uint64_t)std::max(4.0f, std::floor((float)N / 1e6f));
        const uint64_t nb = (N + gpb - 1) / gpb;
        b += nb * sizeof(CullNode) + ((nb + 31) / 32 + 2) * 4;
    }
    const uint64_t groups = (N + GPS_GROUP - 1) / GPS_GROUP;
    if (N > 0) b += groups * sizeof(GroupBox) + round_up(groups, 4) + 4;
    if (N > 0 && !st->fully_disable_hierarchical_culling) {
        const uint64_t gpb = st->gaussians_per_block > 0 ? (uint64_t)st->gaussians_per_block
                                                         : (uint64_t)std::max(4.0f, std::floor((float)N / 1e6f));
        b += (N + gpb - 1) / gpb * 4 + (N + GPS_PRE_THREADS - 1) / GPS_PRE_THREADS;
    }
    return b;
}

uint64_t gps_pass_device_bytes(const gps_pass* p) {
    if (!p) return 0;
    const uint64_t N = p->N, npad = (uint64_t)p->ntiles * GPS_TILE, nwords = (N + 31) / 32 + 2;
    uint64_t b = npad * 4 + 2 * nwords * 4 + std::max<uint64_t>(N, 1) * (sizeof(SplatRecord) + 8) +
                 4ull * 4 * p->ntiles + ((uint64_t)p->budget / GPS_SPLAT_CHUNK + 2) * 4 + sizeof(FrameCounters) +
                 (uint64_t)p->hiz_size * 4 + 2 * p->fb_cap * 8 + p->dbg_indices_cap * 4;
    if (p->hier_available) b += (uint64_t)p->cull.node_count * sizeof(CullNode) + (p->cull.mask_words + 2) * 4ull;
    if (p->ord_hist)
        b += ((uint64_t)p->budget / GPS_SPLAT_CHUNK + 2) * 2 + 4ull * (GPS_ORD_BINS + GPS_ORD_PHASES) +
             p->fb_cap / (256ull * (uint64_t)p->settings.render_pass_count) * 4;
    if (p->gboxes) b += (uint64_t)p->group.count * sizeof(GroupBox) + round_up(p->group.count, 4) + 4;
    if (p->live) b += (uint64_t)p->cull.node_count * 4 + (N + GPS_PRE_THREADS - 1) / GPS_PRE_THREADS;
    b += std::max<uint64_t>(N, 1) * 16 + 3ull * p->list_cap * 4 + 2ull * (p->list_cap / GPS_TILE + 1) * 4 +
         8ull * 4 * GPS_BIN_HIST + 2ull * 4 * p->job_cap + 32;
    return b;
}

void gps_pass_destroy(gps_pass* p) {
    if (!p) return;
    destroy_events(p);
    dfree(p->weights);
    dfree(p->maskA);
    dfree(p->maskB);
    dfree(p->recs);
    dfree(p->tile_sum);
    dfree(p->tile_cnt);
    dfree(p->tile_off);
    dfree(p->tile_voff);
    dfree(p->cidx);
    dfree(p->coff);
    dfree(p->chunk_start);
    dfree(p->ord_bucket);
    dfree(p->ord_hist);
    dfree(p->ord_lim);
    dfree(p->ord_tmax);
    dfree(p->ctr);
    dfree(p->dbg_indices);
    dfree(p->fbA);
    dfree(p->fbB);
    dfree(p->hiz.mip);
    dfree(p->nodes);
    dfree(p->cull_mask);
    dfree(p->gboxes);
    dfree(p->gocc);
    dfree(p->live);
    dfree(p->wg_live);
    dfree(p->erect);
    dfree(p->ew);
    dfree(p->big);
    for (uint32_t** a : {&p->reg_cnt, &p->reg_pts, &p->reg_ecnt, &p->reg_ptot, &p->reg_eoff, &p->reg_poff,
                         &p->reg_cursor, &p->reg_jobs, &p->job_reg, &p->job_p0, &p->list, &p->lw, &p->lpre, &p->tsum,
                         &p->toff, &p->bin_state})
        dfree(*a);
    delete p;
}

extern "C++" {
namespace gps {
void make_frame_consts(const gps_frame_params* fp, int spp, FrameConsts* out) {
    FrameConsts& f = *out;
    const int ss = fp->supersampling_factor;
    hm::M4 V, P;
    std::memcpy(V.m, fp->view, 64);
    std::memcpy(P.m, fp->proj, 64);
    const hm::M4 cam = hm::affine_inverse(V);
    const hm::M4 VP = hm::mul(P, V);
    std::memcpy(f.V, V.m, 64);
    std::memcpy(f.VP, VP.m, 64);
    f.campos[0] = cam.at(3, 0);
    f.campos[1] = cam.at(3, 1);
    f.campos[2] = cam.at(3, 2);
    const float r = float(fp->width) / fp->height;
    const float t_f_y = std::tan(fp->fovy_degrees * 0.01745329251994329576923690768489f * 0.5f);
    const float fovx = 2.0f * std::atan(t_f_y * r);
    const float t_f_x = std::tan(fovx * 0.5f);
    f.tan_fov[0] = t_f_x;
    f.tan_fov[1] = t_f_y;
    f.focal[0] = (float)(fp->width * ss) / (2.0f * t_f_x);
    f.focal[1] = (float)(fp->height * ss) / (2.0f * t_f_y);
    f.dilation = 0.3f * (float)ss * (float)ss;
    f.near_plane = fp->near_plane;
    f.far_plane = fp->far_plane;
    f.up_w = fp->width * ss;
    f.up_h = fp->height * ss;
    f.ss = ss;
    f.ss_shift = (ss > 0 && (ss & (ss - 1)) == 0) ? __builtin_ctz((unsigned)ss) : -1;
    f.spp = spp;
    f.Ki = fp->reduce_point_count ? ss * ss : 1;
    f.K = fp->reduce_point_count ? (float)(ss * ss) : 1.0f;
    f.point_cap = det_f2u_rz((float)(f.up_w * f.up_h) / (2.0f * f.K));
    f.occl = fp->enable_occlusion_culling ? 1 : 0;
    f.hier = fp->enable_hierarchical_culling ? 1 : 0;
    f.cull_small = fp->cull_small_gaussians ? 1 : 0;
    f.unbiased = fp->use_unbiased_2d_splatting ? 1 : 0;
    f.reduce = fp->reduce_point_count ? 1 : 0;
    f.frame_seed = 0;
    f.only_prev = 0;
    f.vn2 = (float)(view_norm2_bound(f.V) * (1.0 + 1e-6));
    f.hiz_max = fp->occlusion_culling_min_reduce_count == 0 ? 1 : 0;
}
}  // namespace gps
}

gps_status gps_pass_render(gps_pass* p, const gps_frame_params* fp, float* d_color, float* d_depth, void* stream) {
    if (!p || !fp || !d_color || !d_depth) return GPS_ERR_INVALID_ARG;
    const int spp = fp->samples_per_pixel > 0 ? fp->samples_per_pixel : p->settings.render_pass_count;
    if (!p->settings.flexible_frame_size) {
        if (fp->width != p->settings.width || fp->height != p->settings.height ||
            spp != p->settings.render_pass_count) {
            g_last_error = "runtime resolution / spp change is not supported (main.cu:82-86)";
            return GPS_ERR_INVALID_ARG;
        }
    } else {
        if (fp->width < 1 || fp->height < 1 || fp->width > p->settings.width || fp->height > p->settings.height ||
            spp > p->settings.render_pass_count) {
            g_last_error = "frame larger than the pass's width x height x spp";
            return GPS_ERR_INVALID_ARG;
        }
        if (fp->enable_occlusion_culling && (fp->width != p->settings.width || fp->height != p->settings.height)) {
            g_last_error = "occlusion culling needs frames of the pass's size (Hi-Z chain)";
            return GPS_ERR_UNSUPPORTED;
        }
    }
    if (fp->supersampling_factor < 1 || !(fp->far_plane > fp->near_plane)) return GPS_ERR_INVALID_ARG;
    const int ss = fp->supersampling_factor;
    if ((int64_t)fp->width * ss >= 65536 || (int64_t)fp->height * ss >= 65536) return GPS_ERR_UNSUPPORTED;
    if (fb_words(fp->width * ss, fp->height * ss, spp) >= (1ull << 31)) return GPS_ERR_UNSUPPORTED;
    hipStream_t st = (hipStream_t)stream;
    gps_status s = ensure_fb(p, ss, fb_words(fp->width * ss, fp->height * ss, spp));
    if (s != GPS_OK) return s;

    FrameConsts f;
    make_frame_consts(fp, spp, &f);
    const bool hier = fp->enable_hierarchical_culling && p->hier_available;
    f.hier = hier ? 1 : 0;
    hm::M4 V, P;
    std::memcpy(V.m, fp->view, 64);
    std::memcpy(P.m, fp->proj, 64);
    const hm::M4 VP = hm::mul(P, V);

    const uint64_t bg = background_word(fp->background[0], fp->background[1], fp->background[2]);
    p->cur_fb = (fp->frame % 2 == 0) ? p->fbA : p->fbB;
    if (p->timing_cap > 0) {
        p->timing_slot = (int)(p->timing_frames % (uint64_t)p->timing_cap);
        for (int k = 0; k < GPS_NUM_STAGES * kOcc; ++k) p->ev_used[p->timing_slot * GPS_NUM_STAGES * kOcc + k] = 0;
        ++p->timing_frames;
    } else {
        p->timing_slot = -1;
    }
    stage_mark(p, 0, 0, 0, st);
    // The clear also resets the per-frame counters (pass results, culled
    // counts, atomics) and keeps the Hi-Z ticket and the mask iteration.
    launch_clear(p->cur_fb, p->fb_words, bg, st, reinterpret_cast<uint32_t*>(p->ctr), 10, &p->ctr->atomics,
                 (uint32_t)((sizeof(FrameCounters) - offsetof(FrameCounters, atomics)) / 4));
    stage_mark(p, 0, 0, 1, st);

    Mat4 vpm;
    std::memcpy(vpm.m, VP.m, 64);
    // Group occlusion test: needs a max-only Hi-Z chain (no min-reduced levels).
    const bool group_test = f.occl && p->gboxes && p->group_culling && fp->occlusion_culling_min_reduce_count == 0 &&
                            camera_finite(f);
    // Blocks that are whole preprocess workgroups: the block cull marks each
    // workgroup live or not and zeroes the culled blocks' mask words.
    const bool live_mode = hier && p->wg_live && p->cull.gpb % GPS_PRE_THREADS == 0;
    auto render_pass = [&](bool first) {
        const int slot = first ? 0 : 1;
        const uint32_t frame_seed = (uint32_t)(4 * fp->frame + (first ? 0 : 2));
        if (hier || group_test) {
            stage_mark(p, 1, slot, 0, st);
            if (hier)
                launch_cull(p->cull, fp->frustum_planes, vpm, p->hiz, f.occl, &p->ctr->culled_blocks[slot],
                            group_test ? p->live : nullptr, &p->ctr->live_blocks[slot], p->maskA, p->maskB,
                            &p->ctr->iteration, live_mode ? p->wg_live : nullptr, p->N, st);
            // skipped groups are counted with the diagnostic counters only
            uint32_t* gcount = (p->count_rejects || p->count_atomics) ? &p->ctr->culled_groups[slot] : nullptr;
            if (group_test && hier)  // the groups of the blocks the block cull kept
                launch_group_cull_live(p->group, p->cull, p->live, &p->ctr->live_blocks[slot], p->N, f, p->hiz,
                                       p->cus * 8, gcount, st);
            else if (group_test)
                launch_group_cull(p->group, f, p->hiz, gcount, st);
            stage_mark(p, 1, slot, 1, st);
        }
        PreprocessArgs a;
        a.scene = p->scene->view;
        a.f = f;
        a.f.frame_seed = frame_seed;
        a.f.only_prev = first ? 0 : 1;
        a.maskA = p->maskA;
        a.maskB = p->maskB;
        a.iteration = &p->ctr->iteration;
        a.weights = p->weights;
        a.recs = p->recs;
        a.tile_sum = p->tile_sum;
        a.tile_cnt = p->tile_cnt;
        a.cull = p->cull;
        a.hiz = p->hiz;
        a.gocc = group_test ? p->gocc : nullptr;
        if (live_mode) a.wg_live = p->wg_live;
        a.rejects = p->count_rejects ? &p->ctr->rejects[0][0] : nullptr;
        stage_mark(p, 2, slot, 0, st);
        launch_preprocess(a, st);
        stage_mark(p, 2, slot, 1, st);
        DistributeArgs da;
        da.tile_sum = p->tile_sum;
        da.tile_cnt = p->tile_cnt;
        da.ntiles = p->ntiles;
        da.tile_off = p->tile_off;
        da.tile_voff = p->tile_voff;
        da.ctr = &p->ctr->pass[slot];
        da.budget = p->budget;
        da.weights = p->weights;
        da.cidx = p->cidx;
        da.coff = p->coff;
        da.chunk_start = p->chunk_start;
        da.iteration = &p->ctr->iteration;
        da.cull = p->cull;
        da.hier = f.hier;
        da.N = p->N;
        stage_mark(p, 3, slot, 0, st);
        launch_distribute(da, st);
        stage_mark(p, 3, slot, 1, st);
        SplatArgs sa;
        sa.recs = p->recs;
        sa.cidx = p->cidx;
        sa.coff = p->coff;
        sa.chunk_start = p->chunk_start;
        sa.ctr = &p->ctr->pass[slot];
        sa.fb = p->cur_fb;
        sa.atomics_count = p->count_atomics ? &p->ctr->atomics : nullptr;
        sa.spp = f.spp;
        sa.K = f.Ki;
        sa.up_w = f.up_w;
        sa.unbiased = f.unbiased;
        sa.frame_seed = frame_seed + 1u;
        sa.grid = p->splat_grid;
        sa.xcds = p->xcds;
        stage_mark(p, 4, slot, 0, st);
        BinGeom bg_geom;
        // the biased splat adds colours per sample and has only the
        // point-parallel form
        const bool binned = f.unbiased && p->splat_mode != GPS_SPLAT_POINT_PARALLEL &&
                            (p->splat_mode == GPS_SPLAT_BINNED || f.Ki * f.spp >= GPS_BIN_AUTO_SAMPLES) &&
                            binned_geometry(f.up_w, f.up_h, f.spp, &bg_geom);
        p->last_binned = binned ? 1 : 0;
        if (binned) {
            BinnedArgs ba{};
            ba.recs = p->recs;
            ba.cidx = p->cidx;
            ba.coff = p->coff;
            ba.ctr = &p->ctr->pass[slot];
            ba.g = bg_geom;
            ba.sc.spp = f.spp;
            ba.sc.K = f.Ki;
            ba.sc.up_w = f.up_w;
            ba.sc.unbiased = f.unbiased;
            ba.sc.frame_seed = frame_seed + 1u;
            ba.sc.xcds = p->xcds;
            ba.background = bg;
            ba.fb = p->cur_fb;
            ba.erect = p->erect;
            ba.ew = p->ew;
            ba.reg_cnt = p->reg_cnt;
            ba.reg_pts = p->reg_pts;
            ba.reg_ecnt = p->reg_ecnt;
            ba.reg_ptot = p->reg_ptot;
            ba.reg_eoff = p->reg_eoff;
            ba.reg_poff = p->reg_poff;
            ba.reg_cursor = p->reg_cursor;
            ba.reg_jobs = p->reg_jobs;
            ba.job_reg = p->job_reg;
            ba.job_p0 = p->job_p0;
            ba.list = p->list;
            ba.lw = p->lw;
            ba.lpre = p->lpre;
            ba.tsum = p->tsum;
            ba.toff = p->toff;
            ba.big = p->big;
            ba.state = p->bin_state;
            ba.list_cap = p->list_cap;
            ba.job_cap = p->job_cap;
            ba.job_points = GPS_BIN_JOB_POINTS;
            ba.cus = p->cus;
            // the automatic choice also tests the pass's density on the device
            ba.dense_points =
                p->splat_mode == GPS_SPLAT_BINNED ? 0xFFFFFFFFu : binned_dense_points(f.up_w, f.up_h, f.Ki);
            ba.atomics_count = p->count_atomics ? &p->ctr->atomics : nullptr;
            launch_binned_splat(ba, st);
            sa.only_if = p->bin_state + 3;  // stood down: the point-parallel splat takes the pass
        } else {
            sa.only_if = nullptr;
        }
        // Dense passes of many-sample points splat point-parallel nearest
        // first and skip Gaussians behind their tiles (the density is decided
        // on the device from the point count; launched when the scene could
        // make the pass dense).
        const bool ordered =
            f.unbiased && p->ord_hist && p->ordered_mode != GPS_ORDERED_OFF &&
            (p->ordered_mode == GPS_ORDERED_ALWAYS ||
             (f.Ki * f.spp >= GPS_BIN_AUTO_SAMPLES &&
              2ull * p->N * (uint64_t)f.Ki > (uint64_t)GPS_ORD_DENSE_SAMPLES * (uint64_t)f.up_w * (uint64_t)f.up_h));
        if (ordered) {
            OrderedArgs oa;
            oa.s = sa;
            oa.up_w = f.up_w;
            oa.up_h = f.up_h;
            oa.forced = p->ordered_mode == GPS_ORDERED_ALWAYS ? 1 : 0;
            oa.pre_test = first ? 0 : 1;
            oa.bucket = p->ord_bucket;
            oa.hist = p->ord_hist;
            oa.lim = p->ord_lim;
            oa.tmax = p->ord_tmax;
            oa.tw64 = fb_tiles_per_row(f.up_w);
            oa.ntiles = (uint32_t)(p->fb_words / (256ull * (uint64_t)f.spp));
            oa.cus = p->cus;
            oa.skipped_count = p->count_rejects ? &p->ctr->ordered_skipped[slot] : nullptr;
            launch_ordered_splat(oa, st);
        } else {
            launch_splat(sa, st);
        }
        stage_mark(p, 4, slot, 1, st);
    };
    render_pass(true);
    if (fp->enable_occlusion_culling) {
        stage_mark(p, 5, 0, 0, st);
        launch_combine(p->cur_fb, d_color, d_depth, fp->width, fp->height, ss, f.spp, fp->near_plane, fp->far_plane, 1,
                       st);
        stage_mark(p, 5, 0, 1, st);
        stage_mark(p, 6, 0, 0, st);
        launch_hiz_build(p->hiz, d_depth, fp->occlusion_culling_min_reduce_count, &p->ctr->hiz_ticket, st);
        stage_mark(p, 6, 0, 1, st);
        render_pass(false);
    }
    stage_mark(p, 5, 1, 0, st);
    launch_combine(p->cur_fb, d_color, d_depth, fp->width, fp->height, ss, f.spp, fp->near_plane, fp->far_plane, 0, st);
    stage_mark(p, 5, 1, 1, st);
    if (fp->enable_occlusion_culling) {
        stage_mark(p, 6, 1, 0, st);
        launch_hiz_build(p->hiz, d_depth, fp->occlusion_culling_min_reduce_count, &p->ctr->hiz_ticket, st);
        stage_mark(p, 6, 1, 1, st);
    }
    p->last_occl = fp->enable_occlusion_culling ? 1 : 0;
    p->last_hier = f.hier;
    p->last_K = f.Ki;
    p->last_spp = f.spp;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return GPS_OK;
}

}  // extern "C"

extern "C++" {
namespace gps {
// gps_pass_render of a pass that exists only for this call: created as
// gps_pass_create would (fully_disable_hierarchical_culling = 1, point budget
// max_points, the frame's size and spp), with every buffer carved out of the
// device workspace `ws` of `ws_bytes` instead of allocated.  What the kernels
// need at zero (pass counters, tile sums, the empty cull mask) is cleared by
// a fill kernel on `st`; the rest is written before it is read.  The point-
// parallel splat is used (it gives the binned form's framebuffer bit for bit).
gps_status render_in_workspace(gps_scene* scene, const gps_frame_params* fp, int64_t max_points, float* d_color,
                               float* d_depth, hipStream_t st, uint8_t* ws, uint64_t ws_bytes, int cus) {
    if (!scene || !fp || !d_color || !d_depth || !ws || max_points < 0) return GPS_ERR_INVALID_ARG;
    if (fp->enable_occlusion_culling || fp->enable_hierarchical_culling) {
        g_last_error = "gps_frame_render_noalloc renders frames without occlusion or block culling";
        return GPS_ERR_UNSUPPORTED;
    }
    const int spp = fp->samples_per_pixel > 0 ? fp->samples_per_pixel : 1;
    const int ss = fp->supersampling_factor;
    if (fp->width < 1 || fp->height < 1 || ss < 1 || !(fp->far_plane > fp->near_plane)) return GPS_ERR_INVALID_ARG;
    if ((int64_t)fp->width * ss >= 65536 || (int64_t)fp->height * ss >= 65536) return GPS_ERR_UNSUPPORTED;
    const uint64_t N = scene->view.N;
    const uint64_t words = fb_words(fp->width * ss, fp->height * ss, spp);
    if (N > GPS_NOALLOC_MAX_GAUSSIANS) {
        g_last_error = "gps_frame_render_noalloc: more Gaussians than GPS_NOALLOC_MAX_GAUSSIANS";
        return GPS_ERR_UNSUPPORTED;
    }
    if (words > GPS_NOALLOC_MAX_FB_WORDS) {
        g_last_error = "gps_frame_render_noalloc: framebuffer larger than GPS_NOALLOC_MAX_FB_WORDS";
        return GPS_ERR_UNSUPPORTED;
    }
    if ((uint64_t)max_points > GPS_NOALLOC_MAX_POINTS) {
        g_last_error = "gps_frame_render_noalloc: point budget above GPS_NOALLOC_MAX_POINTS";
        return GPS_ERR_UNSUPPORTED;
    }
    gps_pass p;
    p.scene = scene;
    p.settings.width = fp->width;
    p.settings.height = fp->height;
    p.settings.render_pass_count = spp;
    p.settings.supersampling_factor = ss;
    p.settings.max_number_of_points = max_points;
    p.settings.fully_disable_hierarchical_culling = 1;
    p.N = (uint32_t)N;
    p.budget = (uint32_t)max_points;
    p.ntiles = (uint32_t)((N + GPS_TILE - 1) / GPS_TILE);
    const uint64_t npad = (uint64_t)p.ntiles * GPS_TILE;
    const uint64_t nwords = (N + 31) / 32 + 2;
    // Layout: the words cleared per call first (one fill launch), then the rest.
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        uint8_t* q = ws + off;
        off += round_up(bytes, 256);
        return q;
    };
    uint8_t* zero_begin = ws;
    p.ctr = reinterpret_cast<FrameCounters*>(take(sizeof(FrameCounters)));
    p.cull_mask = reinterpret_cast<uint32_t*>(take(2 * 4));
    p.tile_sum = reinterpret_cast<uint32_t*>(take(4ull * p.ntiles));
    p.tile_cnt = reinterpret_cast<uint32_t*>(take(4ull * p.ntiles));
    const uint64_t zero_words = off / 4;
    p.tile_off = reinterpret_cast<uint32_t*>(take(4ull * p.ntiles));
    p.tile_voff = reinterpret_cast<uint32_t*>(take(4ull * p.ntiles));
    p.weights = reinterpret_cast<uint32_t*>(take(4 * npad));
    p.maskA = reinterpret_cast<uint32_t*>(take(4 * nwords));
    p.maskB = reinterpret_cast<uint32_t*>(take(4 * nwords));
    p.recs = reinterpret_cast<SplatRecord*>(take(sizeof(SplatRecord) * std::max<uint64_t>(N, 1)));
    p.cidx = reinterpret_cast<uint32_t*>(take(4 * std::max<uint64_t>(N, 1)));
    p.coff = reinterpret_cast<uint32_t*>(take(4 * std::max<uint64_t>(N, 1)));
    p.chunk_start = reinterpret_cast<uint32_t*>(take(4 * ((uint64_t)p.budget / GPS_SPLAT_CHUNK + 2)));
    p.fbA = p.fbB = reinterpret_cast<uint64_t*>(take(8 * words));
    if (off > ws_bytes) {
        g_last_error = "gps_frame_render_noalloc: workspace too small for this frame";
        return GPS_ERR_UNSUPPORTED;
    }
    p.fb_cap = words;
    p.fb_words = words;
    p.fb_ss = ss;
    p.splat_mode = GPS_SPLAT_POINT_PARALLEL;
    p.hier_available = false;
    const uint32_t gpb = (uint32_t)std::max(4.0f, std::floor((float)N / (float)1e6));
    p.cull.gpb = gpb;
    p.cull.node_count = (uint32_t)((N + gpb - 1) / gpb);
    p.cull.mask_words = (p.cull.node_count + 31) / 32;
    p.cull.nodes = nullptr;
    p.cull.mask = p.cull_mask;
    p.xcds = (cus % 8 == 0) ? 8 : 1;
    p.splat_grid = (uint32_t)cus * GPS_SPLAT_WGS_PER_CU;
    p.cus = (uint32_t)cus;
    launch_zero_u32(reinterpret_cast<uint32_t*>(zero_begin), (uint32_t)zero_words, st);
    return gps_pass_render(&p, fp, d_color, d_depth, st);
}
}  // namespace gps
}

extern "C" {

gps_status gps_pass_set_timing(gps_pass* p, int32_t capacity) {
    if (!p || capacity < 0) return GPS_ERR_INVALID_ARG;
    destroy_events(p);
    if (capacity == 0) return GPS_OK;
    const size_t n = (size_t)capacity * GPS_NUM_STAGES * kOcc * 2;
    p->ev.resize(n);
    for (size_t i = 0; i < n; ++i) GPS_HIP(hipEventCreate(&p->ev[i]));
    p->ev_used.assign(n / 2, 0);
    p->timing_cap = capacity;
    return GPS_OK;
}

gps_status gps_pass_timing_summary(gps_pass* p, double out_ms[GPS_NUM_STAGES], uint64_t out_launches[GPS_NUM_STAGES],
                                   uint64_t* out_frames) {
    if (!p || !out_ms || !out_launches) return GPS_ERR_INVALID_ARG;
    for (int k = 0; k < GPS_NUM_STAGES; ++k) {
        out_ms[k] = 0.0;
        out_launches[k] = 0;
    }
    if (p->timing_cap == 0) return GPS_ERR_NOT_READY;
    const uint64_t frames = std::min<uint64_t>(p->timing_frames, (uint64_t)p->timing_cap);
    for (uint64_t slot = 0; slot < frames; ++slot)
        for (int stg = 0; stg < GPS_NUM_STAGES; ++stg)
            for (int occ = 0; occ < kOcc; ++occ) {
                const int i = ev_index((int)slot, stg, occ, 0);
                if (!p->ev_used[i / 2]) continue;
                GPS_HIP(hipEventSynchronize(p->ev[i + 1]));
                float ms = 0.0f;
                GPS_HIP(hipEventElapsedTime(&ms, p->ev[i], p->ev[i + 1]));
                out_ms[stg] += ms;
                out_launches[stg] += 1;
            }
    if (out_frames) *out_frames = frames;
    p->timing_frames = 0;
    std::fill(p->ev_used.begin(), p->ev_used.end(), 0);
    return GPS_OK;
}

gps_status gps_pass_stats(gps_pass* p, gps_stats* out, void* stream) {
    if (!p || !out) return GPS_ERR_INVALID_ARG;
    FrameCounters c;
    GPS_HIP(hipStreamSynchronize((hipStream_t)stream));
    GPS_HIP(hipMemcpy(&c, p->ctr, sizeof(c), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->num_gaussians = p->N;
    out->passes = p->last_occl ? 2 : 1;
    out->K = p->last_K;
    for (int i = 0; i < out->passes; ++i) {
        out->valid_gaussians[i] = c.pass[i].valid;
        out->total_weight[i] = c.pass[i].total;
        out->num_points[i] = c.pass[i].points;
        out->culled_blocks[i] = c.culled_blocks[i];
        out->culled_groups[i] = c.culled_groups[i];
        out->samples += (uint64_t)c.pass[i].points * (uint64_t)p->last_K * (uint64_t)p->last_spp;
        if (c.pass[i].total > c.pass[i].points) out->budget_truncated = 1;
    }
    out->atomics = c.atomics;
    if (p->last_binned) {
        uint32_t bs[8];
        GPS_HIP(hipMemcpy(bs, p->bin_state, sizeof(bs), hipMemcpyDeviceToHost));
        out->splat_binned = 1;
        out->splat_fallback = (int32_t)bs[3];
        out->bin_entries = bs[0];
        out->bin_jobs = bs[1];
        out->bin_big = bs[2];
    }
    if (p->hier_available) {
        out->gaussians_per_block = (int32_t)p->cull.gpb;
        out->num_blocks = p->cull.node_count;
    }
    return GPS_OK;
}

uint64_t gps_pass_debug_size(gps_pass* p, int32_t which) {
    if (!p) return 0;
    FrameCounters c;
    if (hipDeviceSynchronize() != hipSuccess) return 0;
    if (hipMemcpy(&c, p->ctr, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess) return 0;
    const PassCounters& lp = c.pass[p->last_occl ? 1 : 0];
    switch (which) {
        case GPS_DBG_WEIGHTS: return p->N;
        case GPS_DBG_MASK_CURRENT:
        case GPS_DBG_MASK_PREVIOUS: return (p->N + 31) / 32;
        case GPS_DBG_VALID_IDS:
        case GPS_DBG_VALID_OFFSETS: return lp.valid;
        case GPS_DBG_INDICES: return lp.points;
        case GPS_DBG_FRAMEBUFFER: return p->fb_words;
        case GPS_DBG_RECORDS: return 16ull * p->N;
        case GPS_DBG_HIZ: return p->hiz_size;
        case GPS_DBG_CULL_MASK: return p->cull.mask_words;
        case GPS_DBG_REJECTS: return 22;
    }
    return 0;
}

gps_status gps_pass_debug_copy(gps_pass* p, int32_t which, void* dst, uint64_t count, void* stream) {
    if (!p || (!dst && count)) return GPS_ERR_INVALID_ARG;
    GPS_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (count > gps_pass_debug_size(p, which)) return GPS_ERR_INVALID_ARG;
    if (count == 0) return GPS_OK;
    FrameCounters c;
    GPS_HIP(hipMemcpy(&c, p->ctr, sizeof(c), hipMemcpyDeviceToHost));
    const PassCounters& lp = c.pass[p->last_occl ? 1 : 0];
    // The mask written by the last pass is "current" unless that pass produced
    // points, in which case the iteration has since flipped.
    const uint32_t written = lp.points > 0 ? (c.iteration - 1u) : c.iteration;
    const uint32_t* mask_written = (written & 1u) ? p->maskB : p->maskA;
    const uint32_t* mask_other = (written & 1u) ? p->maskA : p->maskB;
    const void* src = nullptr;
    uint64_t esz = 4;
    switch (which) {
        case GPS_DBG_WEIGHTS: src = p->weights; break;
        case GPS_DBG_MASK_CURRENT: src = mask_written; break;
        case GPS_DBG_MASK_PREVIOUS: src = mask_other; break;
        case GPS_DBG_VALID_IDS: src = p->cidx; break;
        case GPS_DBG_VALID_OFFSETS: src = p->coff; break;
        case GPS_DBG_FRAMEBUFFER: src = p->cur_fb; esz = 8; break;
        case GPS_DBG_RECORDS: src = p->recs; break;
        case GPS_DBG_HIZ: src = p->hiz.mip; break;
        case GPS_DBG_CULL_MASK: src = p->cull_mask; break;
        case GPS_DBG_REJECTS: src = &p->ctr->rejects[0][0]; break;
        case GPS_DBG_INDICES: {
            if (p->dbg_indices_cap < count) {
                dfree(p->dbg_indices);
                gps_status s = dalloc(&p->dbg_indices, count, false);
                if (s != GPS_OK) return s;
                p->dbg_indices_cap = count;
            }
            launch_materialize_indices(p->cidx, p->coff, &p->ctr->pass[p->last_occl ? 1 : 0], p->dbg_indices,
                                       (uint32_t)count, nullptr);
            GPS_HIP(hipDeviceSynchronize());
            src = p->dbg_indices;
            break;
        }
        default: return GPS_ERR_INVALID_ARG;
    }
    GPS_HIP(hipMemcpy(dst, src, count * esz, hipMemcpyDeviceToHost));
    if (which == GPS_DBG_WEIGHTS && p->last_hier) {
        // Slots of culled blocks were not written this pass: their weight is 0.
        std::vector<uint32_t> m(p->cull.mask_words);
        GPS_HIP(hipMemcpy(m.data(), p->cull_mask, m.size() * 4, hipMemcpyDeviceToHost));
        uint32_t* w = (uint32_t*)dst;
        for (uint64_t g = 0; g < count; ++g) {
            const uint64_t b = g / p->cull.gpb;
            if ((m[b >> 5] >> (b & 31)) & 1u) w[g] = 0;
        }
    }
    return GPS_OK;
}

// ------------------------------------------------------------------ resolve
gps_status gps_resolve_taa(const float* cur_color, const float* cur_depth, const float* prev_color,
                           const float* prev_depth, const float reprojection[16], float blend_factor, int32_t width,
                           int32_t height, int32_t frame_number, float* out_color, uint8_t* out_rgb8, void* stream) {
    (void)prev_depth;
    if (!cur_color || !cur_depth || !prev_color || !reprojection || !out_color || !out_rgb8 || width <= 0 || height <= 0)
        return GPS_ERR_INVALID_ARG;
    if (out_color == cur_color) return GPS_ERR_INVALID_ARG;
    Mat4 M;
    std::memcpy(M.m, reprojection, 64);
    launch_taa(cur_color, cur_depth, prev_color, M, blend_factor, width, height, frame_number, out_color, out_rgb8,
               (hipStream_t)stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "taa launch");
    return GPS_OK;
}

gps_status gps_resolve_progressive(float* color, float* accum, int32_t width, int32_t height, int32_t num_frames,
                                   uint8_t* out_rgb8, void* stream) {
    if (!color || !accum || !out_rgb8 || width <= 0 || height <= 0 || num_frames < 1) return GPS_ERR_INVALID_ARG;
    launch_progressive(color, accum, width, height, num_frames, out_rgb8, (hipStream_t)stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "progressive launch");
    return GPS_OK;
}

// ------------------------------------------------------------------ camera math
void gps_look_at(const float eye[3], const float center[3], const float up[3], float out_view[16]) {
    const hm::M4 v = hm::look_at({eye[0], eye[1], eye[2]}, {center[0], center[1], center[2]}, {up[0], up[1], up[2]});
    std::memcpy(out_view, v.m, 64);
}
void gps_perspective(float fovy, float aspect, float zn, float zf, float out[16]) {
    const hm::M4 p = hm::perspective(fovy, aspect, zn, zf);
    std::memcpy(out, p.m, 64);
}
void gps_frustum_planes(float fovy, const float view[16], float aspect, float zn, float zf, float out[24]) {
    hm::M4 v;
    std::memcpy(v.m, view, 64);
    hm::frustum_planes(fovy, v, aspect, zn, zf, out);
}
void gps_frame_params_init(gps_frame_params* f, const float eye[3], const float center[3], const float up[3],
                           float fovy, int32_t width, int32_t height, int32_t ss, float zn, float zf) {
    std::memset(f, 0, sizeof(*f));
    gps_look_at(eye, center, up, f->view);
    const float aspect = float(width) / height;
    gps_perspective(fovy, aspect, zn, zf, f->proj);
    gps_frustum_planes(fovy, f->view, aspect, zn, zf, f->frustum_planes);
    f->fovy_degrees = fovy;
    f->near_plane = zn;
    f->far_plane = zf;
    f->width = width;
    f->height = height;
    f->supersampling_factor = ss;
    f->frame = 1;
    f->enable_occlusion_culling = 1;
    f->enable_hierarchical_culling = 1;
    f->use_unbiased_2d_splatting = 1;
    f->reduce_point_count = 1;
}
void gps_taa_reprojection(const float prev_vp[16], const float view[16], const float proj[16], float out[16]) {
    hm::M4 a, v, p;
    std::memcpy(a.m, prev_vp, 64);
    std::memcpy(v.m, view, 64);
    std::memcpy(p.m, proj, 64);
    const hm::M4 r = hm::mul(hm::mul(a, hm::affine_inverse(v)), hm::inverse(p));
    std::memcpy(out, r.m, 64);
}
void gps_mat4_mul(const float a[16], const float b[16], float out[16]) {
    hm::M4 x, y;
    std::memcpy(x.m, a, 64);
    std::memcpy(y.m, b, 64);
    const hm::M4 r = hm::mul(x, y);
    std::memcpy(out, r.m, 64);
}

// ------------------------------------------------------------------ plumbing
gps_status gps_device_count(int32_t* out) {
    int n = 0;
    GPS_HIP(hipGetDeviceCount(&n));
    *out = n;
    return GPS_OK;
}
gps_status gps_set_device(int32_t d) {
    GPS_HIP(hipSetDevice(d));
    return GPS_OK;
}
gps_status gps_device_malloc(void** out, uint64_t bytes) {
    GPS_HIP(hipMalloc(out, bytes ? bytes : 1));
    return GPS_OK;
}
gps_status gps_device_free(void* ptr) {
    GPS_HIP(hipFree(ptr));
    return GPS_OK;
}
gps_status gps_memcpy_h2d(void* dst, const void* src, uint64_t bytes, void* stream) {
    GPS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    GPS_HIP(hipStreamSynchronize((hipStream_t)stream));
    return GPS_OK;
}
gps_status gps_memcpy_d2h(void* dst, const void* src, uint64_t bytes, void* stream) {
    GPS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    GPS_HIP(hipStreamSynchronize((hipStream_t)stream));
    return GPS_OK;
}
gps_status gps_memset_d8(void* dst, int32_t value, uint64_t bytes, void* stream) {
    GPS_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
    return GPS_OK;
}
gps_status gps_stream_create(void** out) {
    hipStream_t s;
    GPS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = (void*)s;
    return GPS_OK;
}
gps_status gps_stream_destroy(void* s) {
    GPS_HIP(hipStreamDestroy((hipStream_t)s));
    return GPS_OK;
}
gps_status gps_stream_synchronize(void* s) {
    GPS_HIP(hipStreamSynchronize((hipStream_t)s));
    return GPS_OK;
}
gps_status gps_device_free_memory(uint64_t* free_b, uint64_t* total_b) {
    size_t f = 0, t = 0;
    GPS_HIP(hipMemGetInfo(&f, &t));
    *free_b = f;
    *total_b = t;
    return GPS_OK;
}
gps_status gps_event_create(void** out) {
    hipEvent_t e;
    GPS_HIP(hipEventCreate(&e));
    *out = (void*)e;
    return GPS_OK;
}
gps_status gps_event_destroy(void* e) {
    GPS_HIP(hipEventDestroy((hipEvent_t)e));
    return GPS_OK;
}
gps_status gps_event_record(void* e, void* s) {
    GPS_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
    return GPS_OK;
}
gps_status gps_event_elapsed_ms(void* a, void* b, float* out) {
    GPS_HIP(hipEventSynchronize((hipEvent_t)b));
    GPS_HIP(hipEventElapsedTime(out, (hipEvent_t)a, (hipEvent_t)b));
    return GPS_OK;
}

}  // extern "C"

// IEEE binary16 round-to-nearest-even from binary32.
uint16_t gps_float_to_half_rne(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u) return (uint16_t)(sign | 0x7C00u | (ax > 0x7F800000u ? 0x200u : 0u));
    if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // rounds to >= 65520 -> inf
    if (ax < 0x38800000u) {                                   // < 2^-14: subnormal half
        if (ax < 0x33000000u) return (uint16_t)sign;          // < 2^-25 rounds to 0
        const uint32_t e = ax >> 23;                          // 102..112
        const uint32_t m = (ax & 0x7FFFFFu) | 0x800000u;
        const uint32_t drop = 126u - e;                       // 14..24
        uint32_t q = m >> drop;
        const uint32_t rem = m & ((1u << drop) - 1u);
        const uint32_t half = 1u << (drop - 1u);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return (uint16_t)(sign | q);
    }
    uint32_t h = ((ax >> 13) - (112u << 10));
    const uint32_t rem = ax & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
}
