// gps_server.cpp — serving loop over resident scenes with continuous batching.
//
// The reference renders one camera per frame from a single host loop
// (RenderServer::renderFrame, src/core/rendering/render_server.cu:86-160).
// This engine keeps several scenes resident and serves independent render
// requests of mixed resolution / spp on one GPU:
//
//   * slots: each slot is a HIP stream (default priority) with its output
//     buffers and pinned staging; a request runs entirely on one slot;
//   * passes: a slot holds one gps_pass per scene it has served (created at
//     the server's maximum frame size with flexible_frame_size = 1), created
//     lazily within the HBM budget and evicted least-recently-used;
//   * scheduler: a host thread that, every tick, retires finished slots
//     (hipEventQuery) and admits every arrived request for which a free slot
//     with (or with budget for) a pass of its scene exists, in arrival order
//     without head-of-line blocking.  Requests that arrive while others are
//     on the GPU are admitted as soon as a slot and memory free up, and the
//     work of different slots overlaps on the device.
//
// A request's frame is stateless: occlusion culling (which needs the
// previous frame of the same camera) is turned off, block frustum culling is
// kept, and the resolve is the progressive resolve of one frame.  The result
// therefore depends only on the request, never on which slot ran it or what
// ran beside it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gps.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    float* color = nullptr;
    float* accum = nullptr;
    float* depth = nullptr;
    uint8_t* rgb = nullptr;
    uint8_t* h_rgb = nullptr;   // pinned staging
    float* h_color = nullptr;   // pinned staging
    std::vector<gps_pass*> pass;       // per scene
    std::vector<uint64_t> pass_used;   // LRU stamp per scene
    bool busy = false;
    uint64_t ticket = 0;
    gps_request req{};
    double arrival = 0, admit = 0;
    double busy_s = 0;
};

struct Pending {
    uint64_t ticket;
    gps_request req;
};

}  // namespace

struct gps_server {
    gps_server_config cfg{};
    std::vector<gps_scene*> scenes;
    std::vector<Slot> slots;
    std::deque<Pending> queue;
    std::vector<gps_request_result> results;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread thread;
    std::atomic<bool> stop{false};
    bool started = false;
    Clock::time_point t0;
    uint64_t next_ticket = 1;
    uint64_t in_flight = 0;  // queued + running
    uint64_t budget = 0, used = 0;
    uint64_t stamp = 0;
    gps_server_stats st{};
    std::string error;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

gps_status hip_status(hipError_t e, gps_server* s, const char* what) {
    if (e == hipSuccess) return GPS_OK;
    s->error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? GPS_ERR_OUT_OF_MEMORY : GPS_ERR_HIP;
}

// Settings of every slot pass: the largest request, any smaller frame.
gps_render_settings slot_settings(const gps_server* s) {
    gps_render_settings rs{};
    rs.width = s->cfg.max_width;
    rs.height = s->cfg.max_height;
    rs.render_pass_count = s->cfg.max_spp;
    rs.supersampling_factor = s->cfg.max_supersampling;
    rs.max_number_of_points = 250000000;
    rs.fully_disable_hierarchical_culling = 0;
    rs.gaussians_per_block = 0;
    rs.flexible_frame_size = 1;
    return rs;
}

// Pass of `scene` on slot `k`, created (and others evicted) within budget.
// Returns nullptr when the budget cannot hold it right now.
gps_pass* slot_pass(gps_server* s, int k, int scene) {
    Slot& sl = s->slots[k];
    if (sl.pass[scene]) {
        sl.pass_used[scene] = ++s->stamp;
        return sl.pass[scene];
    }
    const gps_render_settings rs = slot_settings(s);
    const uint64_t need = gps_pass_estimate_bytes(&rs, s->scenes[scene]);
    while (s->used + need > s->budget) {
        // Evict the least recently used pass held by an idle slot.
        int bk = -1, bs = -1;
        uint64_t best = ~0ull;
        for (int j = 0; j < (int)s->slots.size(); ++j) {
            if (s->slots[j].busy) continue;
            for (int q = 0; q < (int)s->scenes.size(); ++q)
                if (s->slots[j].pass[q] && s->slots[j].pass_used[q] < best) {
                    best = s->slots[j].pass_used[q];
                    bk = j;
                    bs = q;
                }
        }
        if (bk < 0) return nullptr;  // wait until a slot finishes
        s->used -= gps_pass_device_bytes(s->slots[bk].pass[bs]);
        gps_pass_destroy(s->slots[bk].pass[bs]);
        s->slots[bk].pass[bs] = nullptr;
        ++s->st.passes_evicted;
    }
    gps_pass* p = nullptr;
    if (gps_pass_create(&rs, s->scenes[scene], &p) != GPS_OK) {
        s->error = std::string("pass create: ") + gps_last_error();
        return nullptr;
    }
    s->used += gps_pass_device_bytes(p);
    sl.pass[scene] = p;
    sl.pass_used[scene] = ++s->stamp;
    ++s->st.passes_created;
    return p;
}

void finish(gps_server* s, int k, gps_status status) {
    Slot& sl = s->slots[k];
    const gps_frame_params& f = sl.req.frame;
    const size_t px = (size_t)f.width * f.height;
    if (status == GPS_OK) {
        if (sl.req.out_rgb8) std::memcpy(sl.req.out_rgb8, sl.h_rgb, px * 3);
        if (sl.req.out_color) std::memcpy(sl.req.out_color, sl.h_color, px * 12);
    }
    gps_request_result r{};
    r.ticket = sl.ticket;
    r.status = status;
    r.slot = k;
    r.arrival_s = sl.arrival;
    r.admit_s = sl.admit;
    r.done_s = s->now();
    sl.busy_s += r.done_s - sl.admit;
    sl.busy = false;
    std::lock_guard<std::mutex> g(s->mu);
    s->results.push_back(r);
    --s->in_flight;
    ++s->st.completed;
    s->cv_done.notify_all();
}

gps_status launch(gps_server* s, int k, gps_pass* p) {
    Slot& sl = s->slots[k];
    gps_frame_params f = sl.req.frame;
    f.enable_occlusion_culling = 0;
    const size_t px = (size_t)f.width * f.height;
    gps_status e = gps_pass_render(p, &f, sl.color, sl.depth, sl.stream);
    if (e == GPS_OK) e = gps_resolve_progressive(sl.color, sl.accum, f.width, f.height, 1, sl.rgb, sl.stream);
    if (e != GPS_OK) return e;
    // The image always comes back to host memory (what a server sends out).
    if (gps_status h = hip_status(hipMemcpyAsync(sl.h_rgb, sl.rgb, px * 3, hipMemcpyDeviceToHost, sl.stream), s,
                                  "rgb copy"))
        return h;
    if (sl.req.out_color)
        if (gps_status h = hip_status(
                hipMemcpyAsync(sl.h_color, sl.color, px * 12, hipMemcpyDeviceToHost, sl.stream), s, "color copy"))
            return h;
    return hip_status(hipEventRecord(sl.done, sl.stream), s, "event record");
}

void scheduler(gps_server* s) {
    (void)hipSetDevice(s->cfg.device);
    const double tick = std::max(1, s->cfg.tick_us) * 1e-6;
    while (!s->stop.load()) {
        // retire
        for (int k = 0; k < (int)s->slots.size(); ++k) {
            Slot& sl = s->slots[k];
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipSuccess) finish(s, k, GPS_OK);
            else if (q != hipErrorNotReady) finish(s, k, hip_status(q, s, "request"));
        }
        // admit: every arrived request that finds a slot, in arrival order
        const double t = s->now();
        double next_arrival = 1e30;
        std::vector<Pending> admit;
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->st.max_queue = std::max<uint64_t>(s->st.max_queue, s->queue.size());
            for (auto it = s->queue.begin(); it != s->queue.end();) {
                if (it->req.arrival_s > t) {
                    next_arrival = std::min(next_arrival, it->req.arrival_s);
                    ++it;
                    continue;
                }
                admit.push_back(*it);
                it = s->queue.erase(it);
            }
        }
        std::vector<Pending> back;
        int batch = 0;
        for (const Pending& pd : admit) {
            int chosen = -1;
            gps_pass* p = nullptr;
            // prefer an idle slot that already holds a pass of this scene
            for (int k = 0; k < (int)s->slots.size() && chosen < 0; ++k)
                if (!s->slots[k].busy && s->slots[k].pass[pd.req.scene]) chosen = k;
            if (chosen < 0)
                for (int k = 0; k < (int)s->slots.size() && chosen < 0; ++k)
                    if (!s->slots[k].busy) chosen = k;
            if (chosen >= 0) p = slot_pass(s, chosen, pd.req.scene);
            if (!p) {
                back.push_back(pd);
                continue;
            }
            Slot& sl = s->slots[chosen];
            sl.busy = true;
            sl.ticket = pd.ticket;
            sl.req = pd.req;
            sl.arrival = pd.req.arrival_s < 0 ? t : pd.req.arrival_s;
            sl.admit = s->now();
            const gps_status e = launch(s, chosen, p);
            if (e != GPS_OK) finish(s, chosen, e);
            ++batch;
        }
        if (batch) {
            ++s->st.ticks_with_admissions;
            s->st.max_batch = std::max<uint64_t>(s->st.max_batch, batch);
        }
        if (!back.empty()) {
            std::lock_guard<std::mutex> g(s->mu);
            for (auto it = back.rbegin(); it != back.rend(); ++it) s->queue.push_front(*it);
        }
        // sleep until the next tick, or earlier for an arrival / new submit
        std::unique_lock<std::mutex> g(s->mu);
        const double wait = std::min(tick, std::max(0.0, next_arrival - s->now()));
        s->cv_work.wait_for(g, std::chrono::duration<double>(wait));
    }
}

}  // namespace

extern "C" {

gps_status gps_server_create(const gps_server_config* cfg, gps_server** out) {
    if (!cfg || !out) return GPS_ERR_INVALID_ARG;
    *out = nullptr;
    if (cfg->slots < 1 || cfg->slots > 64 || cfg->max_width < 1 || cfg->max_height < 1 || cfg->max_spp < 1 ||
        cfg->max_supersampling < 1)
        return GPS_ERR_INVALID_ARG;
    if (hipSetDevice(cfg->device) != hipSuccess) return GPS_ERR_HIP;
    gps_server* s = new gps_server();
    s->cfg = *cfg;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    s->budget = cfg->hbm_budget_bytes ? cfg->hbm_budget_bytes : (uint64_t)(0.9 * (double)free_b);
    const size_t px = (size_t)cfg->max_width * cfg->max_height;
    s->slots.resize(cfg->slots);
    for (Slot& sl : s->slots) {
        hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.color, px * 12);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.accum, px * 12);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.depth, px * 4);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.rgb, px * 3);
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_rgb, px * 3, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_color, px * 12, hipHostMallocDefault);
        if (e != hipSuccess) {
            gps_server_destroy(s);
            return e == hipErrorOutOfMemory ? GPS_ERR_OUT_OF_MEMORY : GPS_ERR_HIP;
        }
        s->used += px * 31;
    }
    *out = s;
    return GPS_OK;
}

gps_status gps_server_add_scene(gps_server* s, gps_scene* scene, int32_t* id) {
    if (!s || !scene || !id || s->started) return GPS_ERR_INVALID_ARG;
    s->scenes.push_back(scene);
    for (Slot& sl : s->slots) {
        sl.pass.push_back(nullptr);
        sl.pass_used.push_back(0);
    }
    *id = (int32_t)s->scenes.size() - 1;
    return GPS_OK;
}

gps_status gps_server_prepare(gps_server* s) {
    if (!s || s->started) return GPS_ERR_INVALID_ARG;
    for (int q = 0; q < (int)s->scenes.size(); ++q)
        for (int k = 0; k < (int)s->slots.size(); ++k) {
            const gps_render_settings rs = slot_settings(s);
            if (s->slots[k].pass[q] || s->used + gps_pass_estimate_bytes(&rs, s->scenes[q]) > s->budget) continue;
            if (!slot_pass(s, k, q)) return GPS_ERR_OUT_OF_MEMORY;
        }
    return hipDeviceSynchronize() == hipSuccess ? GPS_OK : GPS_ERR_HIP;
}

gps_status gps_server_start(gps_server* s) {
    if (!s || s->started || s->scenes.empty()) return GPS_ERR_INVALID_ARG;
    s->t0 = Clock::now();
    s->started = true;
    s->thread = std::thread(scheduler, s);
    return GPS_OK;
}

double gps_server_clock(const gps_server* s) { return s && s->started ? s->now() : 0.0; }

gps_status gps_server_submit(gps_server* s, const gps_request* r, uint64_t* ticket) {
    if (!s || !r || !s->started) return GPS_ERR_INVALID_ARG;
    if (r->scene < 0 || r->scene >= (int32_t)s->scenes.size()) return GPS_ERR_INVALID_ARG;
    const gps_frame_params& f = r->frame;
    const int spp = f.samples_per_pixel > 0 ? f.samples_per_pixel : s->cfg.max_spp;
    if (f.width < 1 || f.height < 1 || f.width > s->cfg.max_width || f.height > s->cfg.max_height ||
        spp > s->cfg.max_spp || f.supersampling_factor < 1 || f.supersampling_factor > s->cfg.max_supersampling)
        return GPS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    Pending pd{s->next_ticket++, *r};
    if (pd.req.arrival_s < 0) pd.req.arrival_s = s->now();
    // keep the queue ordered by arrival time
    auto it = s->queue.end();
    while (it != s->queue.begin() && std::prev(it)->req.arrival_s > pd.req.arrival_s) --it;
    s->queue.insert(it, pd);
    ++s->in_flight;
    ++s->st.submitted;
    if (ticket) *ticket = pd.ticket;
    s->cv_work.notify_all();
    return GPS_OK;
}

int32_t gps_server_poll(gps_server* s, gps_request_result* out, int32_t max, int32_t wait_ms) {
    if (!s || (!out && max > 0)) return -1;
    std::unique_lock<std::mutex> g(s->mu);
    if (s->results.empty() && wait_ms > 0)
        s->cv_done.wait_for(g, std::chrono::milliseconds(wait_ms), [&] { return !s->results.empty(); });
    const int32_t n = std::min<int32_t>(max, (int32_t)s->results.size());
    for (int32_t i = 0; i < n; ++i) out[i] = s->results[i];
    s->results.erase(s->results.begin(), s->results.begin() + n);
    return n;
}

gps_status gps_server_drain(gps_server* s, int32_t timeout_ms) {
    if (!s || !s->started) return GPS_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> g(s->mu);
    const bool ok = s->cv_done.wait_for(g, std::chrono::milliseconds(timeout_ms), [&] { return s->in_flight == 0; });
    return ok ? GPS_OK : GPS_ERR_NOT_READY;
}

gps_status gps_server_get_stats(gps_server* s, gps_server_stats* out) {
    if (!s || !out) return GPS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->st;
    out->hbm_budget_bytes = s->budget;
    out->hbm_used_bytes = s->used;
    out->clock_s = s->started ? s->now() : 0.0;
    double busy = 0;
    for (const Slot& sl : s->slots) busy += sl.busy_s;
    out->slot_busy_s = busy;
    return GPS_OK;
}

const char* gps_server_last_error(const gps_server* s) { return s ? s->error.c_str() : ""; }

void gps_server_destroy(gps_server* s) {
    if (!s) return;
    if (s->started) {
        s->stop.store(true);
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->cv_work.notify_all();
        }
        s->thread.join();
    }
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        for (gps_pass* p : sl.pass) gps_pass_destroy(p);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        (void)hipFree(sl.color);
        (void)hipFree(sl.accum);
        (void)hipFree(sl.depth);
        (void)hipFree(sl.rgb);
        if (sl.h_rgb) (void)hipHostFree(sl.h_rgb);
        if (sl.h_color) (void)hipHostFree(sl.h_color);
    }
    delete s;
}

}  // extern "C"
