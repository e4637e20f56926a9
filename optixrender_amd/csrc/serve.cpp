// Render server: continuous batching of render requests over resident scenes.
//
// The reference renders one frame per Render() call from the viewer loop
// (main.cpp:66-128); this server is new.  Scenes and their BVHs stay resident
// in the context; each request gets its own accumulation buffers (colour,
// albedo, normal, RGBA8).  Several runners -- threads, each with its own path
// pool and HIP stream -- repeatedly
//   1. admit pending requests (FIFO) while a slot is free and their buffers
//      fit in the HBM budget left over by the runners' wavefront buffers,
//   2. take the oldest active request that no other runner holds, with the
//      unheld active requests that share its (scene, max depth, RR start,
//      camera medium), and hand out frames round-robin up to the tick's path
//      budget (every request gets at least one frame),
//   3. run one shared sample window over all of them on the runner's path
//      pool (extend, shade, connect, regenerate until every sample is done,
//      then accumulate) and tone-map the requests that finished, recording
//      their latency.
// So the ticks of different groups -- requests over other scenes, or the
// next requests of a busy scene -- run at the same time on the device, each
// on its own stream, and a runner picks up new work as soon as its tick
// ends, without waiting for the others.  A path is (pixel, frameID) of its
// own request, so a request's image does not depend on what it shared the
// device with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "internal.h"

using namespace oxr;
using Clock = std::chrono::steady_clock;

namespace {

constexpr int kMaxRunners = 8;
constexpr int kDefaultRunners = 2;

enum State { kPending = 0, kActive = 1, kDone = 2, kCancelled = 3, kFailed = 4 };

struct Request {
    uint64_t id;
    oxr_request r;
    SceneDev* scene;
    State state = kPending;
    int frames_done = 0;
    size_t bytes = 0;
    float4 *color = nullptr, *albedo = nullptr, *normal = nullptr, *denoised = nullptr;
    float* depth = nullptr;
    uint32_t* rgba = nullptr;
    uint8_t* cover = nullptr;  // camera-ray cover mask (one byte per 8x8-pixel tile)
    bool cleared = false;      // accumulation buffers zeroed (by the stream of its first tick)
    bool in_tick = false;      // a runner's tick is rendering it
    Clock::time_point submitted, admitted, finished;
    std::string error;
    // The worker holds a request from submit until it has dropped it from the
    // pending queue, the active list and any tick (done, failed or cancelled
    // and dropped).  A request released by the caller while still held is
    // only marked; the worker frees it when it lets go.
    bool held = true;
    bool released = false;
};

// colour, albedo, normal (float4), depth (float), RGBA8, the denoised colour,
// and the cover mask
size_t request_bytes(const oxr_request& r) {
    return (size_t)r.width * r.height * (3 * 16 + 4 + 4 + (r.denoise ? 16 : 0)) +
           (size_t)cover_tiles(r.width) * cover_tiles(r.height);
}

}  // namespace

// A runner: a thread with its own path pool and stream; its tick is one
// sample window over one group of requests.
struct Runner {
    WaveEngine eng;
    hipStream_t stream = nullptr;
    ReqDesc* reqs_dev = nullptr;
    ReqDesc* reqs_host = nullptr;  // pinned staging
    std::thread thread;
};

struct oxr_server {
    oxr_ctx* ctx = nullptr;
    size_t budget = 0, wave_bytes = 0, in_use = 0;  // wave_bytes: all runners' wavefront buffers
    int max_active = 0;
    int64_t tick_paths = 0, cap = 0, pool = 0;
    int trace_grid = 0, shade_grid = 0, sky_grid = 0, stack_lds = 32, iters_per_poll = 2;
    int64_t tail_paths = 0;
    std::vector<Runner*> runners;
    int max_reqs = 0;
    // requests
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::deque<Request*> pending;
    std::vector<Request*> active;
    std::unordered_map<uint64_t, Request*> all;
    uint64_t next_id = 1;
    bool stop = false;
    // statistics
    uint64_t submitted = 0, completed = 0, cancelled = 0, rejected = 0, ticks = 0, paths = 0;
    uint64_t overlapped = 0;          // ticks that began while another runner's tick ran
    int running = 0, max_running = 0;  // ticks in progress, and the most at once
    double busy_ms = 0.0;
    std::vector<double> latencies;
    Clock::time_point started;
    std::string err;
};

namespace {

template <typename T>
hipError_t salloc(T** p, size_t n) {
    *p = nullptr;
    return n ? hipMalloc((void**)p, n * sizeof(T)) : hipSuccess;
}
template <typename T>
void sfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void free_request_buffers(oxr_server* s, Request* q) {
    sfree(q->color); sfree(q->albedo); sfree(q->normal); sfree(q->rgba); sfree(q->depth); sfree(q->denoised);
    sfree(q->cover);
    if (q->bytes) s->in_use -= q->bytes;
    q->bytes = 0;
}

// The worker lets go of a request (under the lock): it is in no list and no
// tick any more.  Its hold on the scene ends, and a request the caller has
// already released is freed now.
void let_go(oxr_server* s, Request* q) {
    if (!q->held) return;
    q->held = false;
    q->scene->server_refs.fetch_sub(1);
    if (q->released) {
        free_request_buffers(s, q);
        s->all.erase(q->id);
        delete q;
    }
}

// A request the caller may still name (not released).
Request* find_request(oxr_server* s, uint64_t id) {
    auto it = s->all.find(id);
    return it == s->all.end() || it->second->released ? nullptr : it->second;
}

bool same_group(const Request* a, const Request* b) {
    return a->scene == b->scene && a->r.max_depth == b->r.max_depth && a->r.rr_start == b->r.rr_start &&
           a->r.camera.medium == b->r.camera.medium;
}

// Admission, under the lock: FIFO, stop at the first request that does not fit.
void admit(oxr_server* s) {
    while (!s->pending.empty() && (int)s->active.size() < s->max_active) {
        Request* q = s->pending.front();
        if (q->state == kCancelled) { s->pending.pop_front(); let_go(s, q); continue; }
        size_t need = request_bytes(q->r);
        if (s->in_use + need + s->wave_bytes > s->budget) break;
        hipError_t e;
        size_t np = (size_t)q->r.width * q->r.height;
        if ((e = salloc(&q->color, np)) != hipSuccess || (e = salloc(&q->albedo, np)) != hipSuccess ||
            (e = salloc(&q->normal, np)) != hipSuccess || (e = salloc(&q->rgba, np)) != hipSuccess ||
            (e = salloc(&q->depth, np)) != hipSuccess ||
            (e = salloc(&q->cover, (size_t)cover_tiles(q->r.width) * cover_tiles(q->r.height))) != hipSuccess ||
            (q->r.denoise && (e = salloc(&q->denoised, np)) != hipSuccess)) {
            free_request_buffers(s, q);
            q->state = kFailed;
            q->error = std::string("allocation: ") + hipGetErrorString(e);
            s->pending.pop_front();
            s->done_cv.notify_all();
            let_go(s, q);
            continue;
        }
        q->bytes = need;
        s->in_use += need;
        q->state = kActive;
        q->admitted = Clock::now();
        s->pending.pop_front();
        s->active.push_back(q);
    }
}

// One tick on runner rn (without the lock; the group's requests are marked
// in_tick, so nothing else touches them).  Returns the paths it rendered.
hipError_t run_tick(oxr_server* s, Runner& rn, std::vector<Request*>& group, std::vector<int>& frames,
                    int64_t& paths) {
    oxr_ctx* c = s->ctx;
    SceneDev* scene = group[0]->scene;
    const oxr_request& r0 = group[0]->r;
    int64_t path = 0, pix = 0;
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < group.size(); i++) {
        Request* q = group[i];
        if (!q->cleared) {  // first tick of the request: its running means start at zero
            const size_t np = (size_t)q->r.width * q->r.height;
            if ((e = hipMemsetAsync(q->depth, 0, 4 * np, rn.stream)) != hipSuccess ||
                (e = hipMemsetAsync(q->color, 0, 16 * np, rn.stream)) != hipSuccess ||
                (e = hipMemsetAsync(q->albedo, 0, 16 * np, rn.stream)) != hipSuccess ||
                (e = hipMemsetAsync(q->normal, 0, 16 * np, rn.stream)) != hipSuccess)
                return e;
            q->cleared = true;
        }
        ReqDesc& d = rn.reqs_host[i];
        std::memset(&d, 0, sizeof(d));
        d.path_begin = path;
        d.pix_begin = pix;
        d.width = q->r.width;
        d.height = q->r.height;
        d.frame0 = q->frames_done;
        d.frames = frames[i];
        d.filter = q->r.filter;
        CamBasis cb = camera_basis(q->r.camera, q->r.width, q->r.height);
        std::memcpy(d.cam_pos, cb.pos, 12);
        std::memcpy(d.cam_dir, cb.dir, 12);
        std::memcpy(d.cam_h, cb.h, 12);
        std::memcpy(d.cam_v, cb.v, 12);
        d.color = q->color;
        d.albedo = q->albedo;
        d.normal = q->normal;
        d.depth = q->depth;
        d.cover = q->cover;
        d.tiles_w = cover_tiles(q->r.width);
        d.by_pixels = make_divider((uint32_t)q->r.width * (uint32_t)q->r.height);
        d.by_width = make_divider((uint32_t)q->r.width);
        const int64_t np = (int64_t)q->r.width * q->r.height;
        path += np * frames[i];
        pix += np;
    }
    e = hipMemcpyAsync(rn.reqs_dev, rn.reqs_host, sizeof(ReqDesc) * group.size(), hipMemcpyHostToDevice, rn.stream);
    if (e != hipSuccess) return e;
    WaveParams p;
    fill_scene_params(p, scene, c->lut_bsdf, c->lut_albedo);
    wave_bind(rn.eng, std::min(s->pool, wave_round_pool(path)), p);
    p.cam_medium = r0.camera.medium;
    p.max_depth = r0.max_depth > 0 ? r0.max_depth : 25;
    p.rr_start = r0.rr_start;
    p.reqs = rn.reqs_dev;
    p.nreq = (int)group.size();
    p.stack_lds = s->stack_lds;
    p.resolve = wave_resolve_mode(p, false);
    p.tail_paths = s->tail_paths;
    WindowRun run;
    run.trace_grid = s->trace_grid;
    run.shade_grid = s->shade_grid;
    run.sky_grid = s->sky_grid;
    run.iters_per_poll = s->iters_per_poll;
    run.total_samples = path;
    run.npix_total = pix;
    if ((e = wave_run_window(rn.eng, p, run, rn.stream)) != hipSuccess) return e;
    for (size_t i = 0; i < group.size(); i++) {
        Request* q = group[i];
        if (q->frames_done + frames[i] < q->r.spp) continue;
        const int64_t np = (int64_t)q->r.width * q->r.height;
        if (q->r.denoise) {
            if ((e = denoise_run(c->dn, q->color, q->albedo, q->normal, q->depth, q->r.width, q->r.height, c->num_cu,
                                 q->denoised, rn.stream)) != hipSuccess ||
                (e = tonemap(q->denoised, q->rgba, np, rn.stream)) != hipSuccess)
                return e;
        } else if ((e = tonemap(q->color, q->rgba, np, rn.stream)) != hipSuccess) {
            return e;
        }
    }
    paths = path;
    return hipStreamSynchronize(rn.stream);
}

void runner_loop(oxr_server* s, Runner* rn) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        admit(s);
        // drop cancelled active requests that no tick is rendering
        for (auto it = s->active.begin(); it != s->active.end();) {
            if ((*it)->state == kCancelled && !(*it)->in_tick) {
                Request* q = *it;
                free_request_buffers(s, q);
                it = s->active.erase(it);
                let_go(s, q);
            } else {
                ++it;
            }
        }
        // the group of the oldest active request no tick is rendering
        Request* lead = nullptr;
        for (Request* q : s->active)
            if (!q->in_tick && q->state == kActive) { lead = q; break; }
        if (!lead) {
            s->cv.wait_for(lk, std::chrono::milliseconds(50));
            continue;
        }
        std::vector<Request*> group;
        for (Request* q : s->active)
            if (!q->in_tick && q->state == kActive && same_group(q, lead) && (int)group.size() < s->max_reqs)
                group.push_back(q);
        // Frames round-robin within the tick's path budget; only the oldest
        // request may overshoot it (by at most one of its frames, which the
        // wavefront capacity covers).  Requests left without frames sit out.
        std::vector<int> frames(group.size(), 0);
        int64_t budget = s->tick_paths, total = 0;
        bool progress = true;
        while (progress) {
            progress = false;
            for (size_t i = 0; i < group.size(); i++) {
                const int64_t np = (int64_t)group[i]->r.width * group[i]->r.height;
                int left = group[i]->r.spp - group[i]->frames_done - frames[i];
                if (left <= 0) continue;
                bool first_of_oldest = (i == 0 && frames[0] == 0);
                if (np > budget && !first_of_oldest) continue;
                frames[i]++;
                budget -= np;
                total += np;
                progress = true;
            }
        }
        {
            std::vector<Request*> g2;
            std::vector<int> f2;
            for (size_t i = 0; i < group.size(); i++)
                if (frames[i] > 0) { g2.push_back(group[i]); f2.push_back(frames[i]); }
            group.swap(g2);
            frames.swap(f2);
        }
        if (total > s->cap || group.empty()) {  // cannot happen: submit bounds a frame by cap - tick_paths
            for (Request* q : group) { q->state = kFailed; q->error = "tick exceeds wavefront capacity"; }
            for (Request* q : group) {
                s->active.erase(std::find(s->active.begin(), s->active.end(), q));
                let_go(s, q);
            }
            s->done_cv.notify_all();
            continue;
        }
        for (Request* q : group) q->in_tick = true;
        if (s->running > 0) s->overlapped++;
        s->running++;
        s->max_running = std::max(s->max_running, s->running);
        lk.unlock();
        auto t0 = Clock::now();
        int64_t paths = 0;
        hipError_t e = run_tick(s, *rn, group, frames, paths);
        auto t1 = Clock::now();
        lk.lock();
        s->running--;
        s->ticks++;
        s->paths += (uint64_t)paths;
        s->busy_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
        for (size_t i = 0; i < group.size(); i++) {
            Request* q = group[i];
            q->in_tick = false;
            if (e != hipSuccess) {
                q->state = kFailed;
                q->error = std::string("render tick: ") + hipGetErrorString(e);
            } else {
                q->frames_done += frames[i];
                if (q->frames_done < q->r.spp || q->state == kCancelled) continue;
                q->state = kDone;
                q->finished = t1;
                s->completed++;
                s->latencies.push_back(std::chrono::duration<double, std::milli>(q->finished - q->submitted).count());
            }
            s->active.erase(std::find(s->active.begin(), s->active.end(), q));
            let_go(s, q);
        }
        s->done_cv.notify_all();
        s->cv.notify_all();  // requests of this tick that want more frames are free for any runner
    }
}

}  // namespace

extern "C" {

int oxr_server_create(oxr_ctx* c, const oxr_server_config* cfg, oxr_server** out) {
    if (!c || !cfg || !out) return OXR_ERR_INVALID;
    *out = nullptr;
    if (cfg->max_active < 1 || cfg->tick_paths < 1 || cfg->max_active > 256 || cfg->max_frame_pixels < 1 ||
        cfg->max_frame_pixels > (1LL << 30) || cfg->runners < 0 || cfg->runners > kMaxRunners) {
        c->err = "server: max_active in [1, 256], tick_paths >= 1, max_frame_pixels in [1, 2^30] and runners in "
                 "[0, 8] required";
        return OXR_ERR_INVALID;
    }
    if (hipSetDevice(c->device) != hipSuccess) return OXR_ERR_HIP;
    oxr_server* s = new oxr_server();
    s->ctx = c;
    s->budget = cfg->hbm_budget_bytes;
    s->max_active = cfg->max_active;
    s->max_reqs = cfg->max_active;
    s->tick_paths = cfg->tick_paths;
    s->cap = cfg->tick_paths + (int64_t)cfg->max_frame_pixels;  // one frame of the largest request may overshoot
    s->trace_grid = c->trace_grid;
    s->shade_grid = c->shade_grid;
    s->sky_grid = c->sky_grid;
    s->stack_lds = c->tuning.stack_lds;
    s->tail_paths = c->tuning.tail_paths;
    s->iters_per_poll = c->tuning.iters_per_poll;
    s->pool = wave_round_pool(std::min<int64_t>(c->tuning.pool_paths, s->cap));
    const int nr = cfg->runners > 0 ? cfg->runners : kDefaultRunners;
    s->wave_bytes = (size_t)nr * wave_bytes(s->pool, s->cap, s->trace_grid);
    auto bail = [&](int code, const char* m) {
        c->err = m;
        oxr_server_destroy(s);
        return code;
    };
    if (s->wave_bytes >= s->budget) return bail(OXR_ERR_INVALID, "server: budget smaller than the wavefront buffers");
    if (s->cap > ((int64_t)1 << 31) - 1) return bail(OXR_ERR_INVALID, "server: tick larger than 2^31 samples");
    for (int i = 0; i < nr; i++) {
        Runner* rn = new Runner();
        s->runners.push_back(rn);
        if (hipStreamCreateWithFlags(&rn->stream, hipStreamNonBlocking) != hipSuccess)
            return bail(OXR_ERR_HIP, "server: stream");
        if (wave_reserve(rn->eng, s->pool, s->cap, s->trace_grid) != hipSuccess || salloc(&rn->reqs_dev, s->max_reqs))
            return bail(OXR_ERR_OOM, "server: wavefront allocation failed");
        if (hipHostMalloc((void**)&rn->reqs_host, sizeof(ReqDesc) * s->max_reqs, 0) != hipSuccess)
            return bail(OXR_ERR_OOM, "server: pinned staging");
    }
    s->started = Clock::now();
    for (Runner* rn : s->runners) rn->thread = std::thread(runner_loop, s, rn);
    *out = s;
    return OXR_OK;
}

int oxr_server_destroy(oxr_server* s) {
    if (!s) return OXR_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    for (Runner* rn : s->runners)
        if (rn->thread.joinable()) rn->thread.join();
    (void)hipSetDevice(s->ctx->device);
    for (Runner* rn : s->runners)
        if (rn->stream) (void)hipStreamSynchronize(rn->stream);
    for (auto& kv : s->all) {
        if (kv.second->held) kv.second->scene->server_refs.fetch_sub(1);
        free_request_buffers(s, kv.second);
        delete kv.second;
    }
    for (Runner* rn : s->runners) {
        wave_release(rn->eng);
        sfree(rn->reqs_dev);
        if (rn->reqs_host) (void)hipHostFree(rn->reqs_host);
        if (rn->stream) (void)hipStreamDestroy(rn->stream);
        delete rn;
    }
    delete s;
    return OXR_OK;
}

const char* oxr_server_last_error(const oxr_server* s) { return s ? s->err.c_str() : "null server"; }

int oxr_submit(oxr_server* s, const oxr_request* r, uint64_t* id) {
    if (!s || !r || !id) return OXR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    std::lock_guard<std::mutex> slk(s->ctx->scenes_mu);  // the context thread may upload or release scenes
    auto it = s->ctx->scenes.find(r->scene_id);
    if (it == s->ctx->scenes.end() || !it->second->built) {
        s->err = "submit: unknown or unbuilt scene";
        s->rejected++;
        return OXR_ERR_INVALID;
    }
    if (r->denoise != 0 && (r->denoise != 1 || !s->ctx->dn_loaded)) {
        s->err = "submit: denoise needs 0 or 1 and the context's denoiser weights (oxr_denoiser_load)";
        s->rejected++;
        return OXR_ERR_INVALID;
    }
    if (r->width <= 0 || r->height <= 0 || r->spp < 1 || r->max_depth < 0 || r->filter < 0 || r->filter > 3 ||
        (int64_t)r->width * r->height > s->cap - s->tick_paths || r->camera.medium < -1 || r->camera.medium >= it->second->n_media) {
        s->err = "submit: invalid request (size, spp, depth, filter or camera medium)";
        s->rejected++;
        return OXR_ERR_INVALID;
    }
    if (request_bytes(*r) + s->wave_bytes > s->budget) {
        s->err = "submit: request can never fit in the HBM budget";
        s->rejected++;
        return OXR_ERR_OOM;
    }
    Request* q = new Request();
    q->id = s->next_id++;
    q->r = *r;
    q->scene = it->second;
    q->scene->server_refs.fetch_add(1);  // oxr_scene_release refuses the scene until the worker lets go
    q->submitted = Clock::now();
    s->pending.push_back(q);
    s->all[q->id] = q;
    s->submitted++;
    *id = q->id;
    s->cv.notify_all();
    return OXR_OK;
}

int oxr_poll(oxr_server* s, uint64_t id, int* state) {
    if (!s || !state) return OXR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    Request* q = find_request(s, id);
    if (!q) return OXR_ERR_INVALID;
    *state = q->state;
    return OXR_OK;
}

int oxr_wait(oxr_server* s, uint64_t id, int timeout_ms) {
    if (!s) return OXR_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    Request* q = find_request(s, id);
    if (!q) return OXR_ERR_INVALID;
    // q stays allocated while this thread waits: only oxr_release_request,
    // which a caller does not overlap with a wait on the same id, frees it
    auto pred = [&] { return q->state == kDone || q->state == kCancelled || q->state == kFailed; };
    if (timeout_ms < 0) s->done_cv.wait(lk, pred);
    else if (!s->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred)) return OXR_ERR_STATE;
    if (q->state == kFailed) {
        s->err = q->error;
        return OXR_ERR_HIP;
    }
    return q->state == kDone ? OXR_OK : OXR_ERR_STATE;
}

int oxr_cancel(oxr_server* s, uint64_t id) {
    if (!s) return OXR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    Request* q = find_request(s, id);
    if (!q) return OXR_ERR_INVALID;
    if (q->state == kPending || q->state == kActive) {
        q->state = kCancelled;
        s->cancelled++;
        s->done_cv.notify_all();
    }
    return OXR_OK;
}

static int copy_result(oxr_server* s, uint64_t id, void* out, size_t bytes_per_px, size_t count, int which) {
    std::lock_guard<std::mutex> lk(s->mu);
    Request* q = find_request(s, id);
    if (!q || !out) return OXR_ERR_INVALID;
    if (q->state != kDone) return OXR_ERR_STATE;
    if (count != (size_t)q->r.width * q->r.height) return OXR_ERR_INVALID;
    const void* src = which == 0 ? (const void*)q->rgba : (q->r.denoise ? (const void*)q->denoised : (const void*)q->color);
    if (hipMemcpy(out, src, bytes_per_px * count, hipMemcpyDeviceToHost) != hipSuccess) return OXR_ERR_HIP;
    return OXR_OK;
}

int oxr_result_rgba8(oxr_server* s, uint64_t id, uint32_t* out, size_t count) {
    return s ? copy_result(s, id, out, 4, count, 0) : OXR_ERR_INVALID;
}
int oxr_result_color(oxr_server* s, uint64_t id, float* out, size_t count) {
    return s ? copy_result(s, id, out, 16, count, 1) : OXR_ERR_INVALID;
}

int oxr_release_request(oxr_server* s, uint64_t id) {
    if (!s) return OXR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    Request* q = find_request(s, id);
    if (!q) return OXR_ERR_INVALID;
    if (q->state == kPending || q->state == kActive) return OXR_ERR_STATE;
    if (q->held) {
        // cancelled, but still in the pending queue, the active list or the
        // tick that is running: the worker frees it when it drops it
        q->released = true;
        s->cv.notify_all();
        return OXR_OK;
    }
    free_request_buffers(s, q);
    s->all.erase(q->id);
    delete q;
    s->cv.notify_all();
    return OXR_OK;
}

int oxr_request_times(oxr_server* s, uint64_t id, double* queued_ms, double* total_ms) {
    if (!s || !queued_ms || !total_ms) return OXR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    Request* q = find_request(s, id);
    if (!q) return OXR_ERR_INVALID;
    if (q->state != kDone) return OXR_ERR_STATE;
    *queued_ms = std::chrono::duration<double, std::milli>(q->admitted - q->submitted).count();
    *total_ms = std::chrono::duration<double, std::milli>(q->finished - q->submitted).count();
    return OXR_OK;
}

int oxr_server_get_stats(oxr_server* s, oxr_server_stats* out) {
    if (!s || !out) return OXR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->cancelled = s->cancelled;
    out->rejected = s->rejected;
    out->ticks = s->ticks;
    out->paths = s->paths;
    out->busy_ms = s->busy_ms;
    out->wall_ms = std::chrono::duration<double, std::milli>(Clock::now() - s->started).count();
    out->active = (int32_t)s->active.size();
    out->pending = (int32_t)s->pending.size();
    out->bytes_in_use = s->in_use + s->wave_bytes;
    out->budget_bytes = s->budget;
    out->overlapped_ticks = s->overlapped;
    out->runners = (int32_t)s->runners.size();
    out->max_running = s->max_running;
    if (!s->latencies.empty()) {
        std::vector<double> l = s->latencies;
        std::sort(l.begin(), l.end());
        auto pct = [&](double q) { return l[std::min(l.size() - 1, (size_t)(q * (l.size() - 1) + 0.5))]; };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        double sum = 0.0;
        for (double v : l) sum += v;
        out->mean_ms = sum / l.size();
    }
    return OXR_OK;
}

}  // extern "C"
