// Host side of the wavefront integrator, shared by the renderer (api.cpp)
// and the request server (serve.cpp): owns the path pool, the shard queues,
// the per-sample outputs of a window and the captured iteration graphs, and
// drives one window from its first sample to the accumulation.
//
// The number of iterations a window needs is data dependent (Russian
// roulette, escapes), so the host replays a graph of `iters_per_poll`
// iterations and, after each, copies the live-path counts (the queue the
// next iteration would trace, the paths it would finish) into pinned memory.  Polls run one graph
// behind the device: the check of replay i overlaps replay i + 1, and the
// window ends at the first replay that leaves no live path (a following
// replay, already queued, finds empty queues and returns at once).
#include <algorithm>
#include <cstring>

#include "internal.h"

namespace oxr {
namespace {

template <typename T>
hipError_t walloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}
template <typename T>
void wfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

constexpr size_t kPollInts = (size_t)kPollIds * kCtrStride;

}  // namespace

// List capacity of one shard: shard k draws the window's 64-sample blocks
// k, k + kShards, ..., at most window / kShards + 64 samples.
int64_t wave_shard_cap(int64_t window) { return window / kShards + 64; }

int64_t wave_round_pool(int64_t paths) {
    const int64_t unit = (int64_t)kShards * 64;
    return std::max<int64_t>(unit, (paths + unit - 1) / unit * unit);
}

size_t wave_bytes(int64_t pool, int64_t window, int trace_grid) {
    // per slot: state 64, hits 2x16, queues 2x2x4 + new-path sample ids 2x4, pending-finish lists 2x4,
    // shadow records 2x48, shadow lists 2x2x4, shade classes 10x4; per sample: outputs 52, sample list 4
    return (size_t)pool * (64 + 32 + 16 + 8 + 8 + 96 + 16 + 40) + (size_t)window * 56 + (size_t)kShards * 64 * 4 +
           gstack_words(trace_grid) * 4 + ctr_ints() * 4 + 4096;
}

void wave_drop_graphs(WaveEngine& e) {
    for (auto& g : e.graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    e.graphs.clear();
}

void wave_release(WaveEngine& e) {
    wave_drop_graphs(e);
    wfree(e.path); wfree(e.hits); wfree(e.sh_env); wfree(e.sh_light); wfree(e.cls_queue); wfree(e.g0); wfree(e.g1);
    wfree(e.s_rad); wfree(e.s_alb); wfree(e.s_nrm); wfree(e.s_dep); wfree(e.q0); wfree(e.q1); wfree(e.gstack);
    wfree(e.samples);
    wfree(e.t0); wfree(e.t1); wfree(e.sl0); wfree(e.sl1);
    wfree(e.ctr); wfree(e.overflow); wfree(e.trav);
    if (e.host_counts) (void)hipHostFree(e.host_counts);
    e.host_counts = nullptr;
    for (auto& ev : e.poll)
        if (ev) (void)hipEventDestroy(ev), ev = nullptr;
    if (e.side_start) (void)hipEventDestroy(e.side_start), e.side_start = nullptr;
    if (e.side_done) (void)hipEventDestroy(e.side_done), e.side_done = nullptr;
    if (e.side) (void)hipStreamDestroy(e.side), e.side = nullptr;
    for (auto ev : e.stage_events) (void)hipEventDestroy(ev);
    e.stage_events.clear();
    e.pool = e.window = 0;
    e.grid = 0;
}

hipError_t wave_reserve(WaveEngine& e, int64_t pool, int64_t window, int trace_grid) {
    hipError_t r;
    if (!e.ctr) {
        if ((r = walloc(&e.ctr, ctr_ints())) != hipSuccess || (r = walloc(&e.overflow, 4)) != hipSuccess ||
            (r = walloc(&e.trav, kTravCounters)) != hipSuccess)
            return r;
        if ((r = hipMemset(e.ctr, 0, ctr_ints() * sizeof(int))) != hipSuccess ||
            (r = hipMemset(e.overflow, 0, 4 * sizeof(int))) != hipSuccess ||
            (r = hipMemset(e.trav, 0, kTravCounters * sizeof(unsigned long long))) != hipSuccess)
            return r;
        if ((r = hipHostMalloc((void**)&e.host_counts, 2 * kPollInts * sizeof(int), 0)) != hipSuccess) return r;
        for (auto& ev : e.poll)
            if ((r = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return r;
        if ((r = hipStreamCreateWithFlags(&e.side, hipStreamNonBlocking)) != hipSuccess ||
            (r = hipEventCreateWithFlags(&e.side_start, hipEventDisableTiming)) != hipSuccess ||
            (r = hipEventCreateWithFlags(&e.side_done, hipEventDisableTiming)) != hipSuccess)
            return r;
    }
    if (pool > e.pool) {
        wave_drop_graphs(e);
        wfree(e.path); wfree(e.hits); wfree(e.sh_env); wfree(e.sh_light); wfree(e.q0); wfree(e.q1);
        wfree(e.t0); wfree(e.t1); wfree(e.sl0); wfree(e.sl1); wfree(e.cls_queue); wfree(e.g0); wfree(e.g1);
        e.pool = 0;
        if ((r = walloc(&e.path, 4 * pool)) != hipSuccess ||
            (r = walloc(&e.hits, 2 * pool)) != hipSuccess || (r = walloc(&e.sh_env, 3 * pool)) != hipSuccess ||
            (r = walloc(&e.sh_light, 3 * pool)) != hipSuccess ||
            (r = walloc(&e.q0, (size_t)kQueueRegions * pool)) != hipSuccess ||
            (r = walloc(&e.q1, (size_t)kQueueRegions * pool)) != hipSuccess || (r = walloc(&e.t0, pool)) != hipSuccess ||
            (r = walloc(&e.t1, pool)) != hipSuccess || (r = walloc(&e.sl0, 2 * pool)) != hipSuccess ||
            (r = walloc(&e.sl1, 2 * pool)) != hipSuccess ||
            (r = walloc(&e.cls_queue, (size_t)kShadeClasses * pool)) != hipSuccess ||
            (r = walloc(&e.g0, pool)) != hipSuccess || (r = walloc(&e.g1, pool)) != hipSuccess)
            return r;
        e.pool = pool;
    }
    if (window > e.window) {
        wave_drop_graphs(e);
        wfree(e.s_rad); wfree(e.s_alb); wfree(e.s_nrm); wfree(e.s_dep); wfree(e.samples);
        e.window = 0;
        if ((r = walloc(&e.s_rad, window)) != hipSuccess || (r = walloc(&e.s_alb, window)) != hipSuccess ||
            (r = walloc(&e.s_nrm, window)) != hipSuccess || (r = walloc(&e.s_dep, window)) != hipSuccess ||
            (r = walloc(&e.samples, (size_t)kShards * wave_shard_cap(window))) != hipSuccess)
            return r;
        e.window = window;
    }
    if (trace_grid > e.grid) {
        wave_drop_graphs(e);
        wfree(e.gstack);
        e.grid = 0;
        if ((r = walloc(&e.gstack, gstack_words(trace_grid))) != hipSuccess) return r;
        e.grid = trace_grid;
    }
    return hipSuccess;
}

void wave_bind(const WaveEngine& e, int64_t pool, WaveParams& p) {
    p.pool = pool;
    p.region = (int)(pool / kShards);
    p.path = e.path;
    p.queue[0] = e.q0; p.queue[1] = e.q1;
    p.fresh_g[0] = e.g0; p.fresh_g[1] = e.g1;
    p.hits = e.hits; p.sh_env = e.sh_env; p.sh_light = e.sh_light; p.cls_queue = e.cls_queue;
    p.term[0] = e.t0; p.term[1] = e.t1;
    p.sh_list[0] = e.sl0; p.sh_list[1] = e.sl1;
    p.s_rad = e.s_rad; p.s_alb = e.s_alb; p.s_nrm = e.s_nrm; p.s_dep = e.s_dep;
    p.samples = e.samples;
    p.shard_cap = wave_shard_cap(e.window);
    p.ctr = e.ctr; p.traversal = e.trav; p.gstack = e.gstack; p.overflow = e.overflow;
}

int wave_grid(int blocks) {
    // every shard needs at least one wave: a multiple of kShards / 4 blocks
    const int unit = kShards / (256 / 64);
    return std::max(unit, (blocks + unit - 1) / unit * unit);
}

hipError_t wave_run_window(WaveEngine& e, const WaveParams& p, WindowRun& r, hipStream_t s) {
    hipError_t err;
    const int K = std::max(2, r.iters_per_poll & ~1);  // even: a replay returns to queue parity 0
    const bool timing = r.stage_timing;
    const bool eager = r.eager || timing;
    r.launches = 0;
    r.iterations = 0;
    // stage ids: 0 regen (window start included), 1 trace (extension and shadow rays), 2 shade,
    // 4 accumulate.
    // With stage timing an event follows every stage; all are read once the
    // window has finished, so timing does not stall the pipelined polls.
    std::vector<int> stages;
    size_t ev_used = 0;
    auto mark = [&](int stage) -> hipError_t {
        if (!timing) return hipSuccess;
        if (ev_used + 1 >= e.stage_events.size()) {
            hipEvent_t ev;
            hipError_t x = hipEventCreate(&ev);
            if (x != hipSuccess) return x;
            e.stage_events.push_back(ev);
        }
        stages.push_back(stage);
        return hipEventRecord(e.stage_events[++ev_used], s);
    };
    auto collect = [&]() -> hipError_t {
        if (!timing) return hipSuccess;
        hipError_t x = hipEventSynchronize(e.stage_events[ev_used]);
        if (x != hipSuccess) return x;
        for (size_t i = 0; i < stages.size(); i++) {
            float ms = 0.0f;
            if ((x = hipEventElapsedTime(&ms, e.stage_events[i], e.stage_events[i + 1])) != hipSuccess) return x;
            r.stage_ms[stages[i]] += ms;
        }
        return hipSuccess;
    };
    auto iterations = [&]() -> hipError_t {
        hipError_t x;
        for (int it = 0; it < K; it++) {
            const int y = it & 1;
            if ((x = wave_trace(p, y, r.trace_grid, r.count, s)) != hipSuccess || (x = mark(1)) != hipSuccess) return x;
            if ((x = wave_regen(p, y, r.shade_grid, s)) != hipSuccess || (x = mark(0)) != hipSuccess) return x;
            if ((x = wave_shade(p, y, r.shade_grid, r.trace_grid, r.count, s)) != hipSuccess || (x = mark(2)) != hipSuccess)
                return x;
        }
        return hipSuccess;
    };
    const int per_iter = 2 + __builtin_popcount(p.class_mask) + (p.tail_paths > 0 ? 1 : 0);

    if (timing) {
        if (e.stage_events.empty()) {
            hipEvent_t ev;
            if ((err = hipEventCreate(&ev)) != hipSuccess) return err;
            e.stage_events.push_back(ev);
        }
        if ((err = hipEventRecord(e.stage_events[0], s)) != hipSuccess) return err;
    }
    if ((err = wave_window_init(p, r.set_req0, r.frame0, r.path_frame_offset, r.frames, r.total_samples, s)) !=
            hipSuccess ||
        (err = build_cover_masks(p.reqs, p.nreq, 0, p.cover_boxes, p.cover_count, s)) != hipSuccess ||
        (err = wave_sky_list(p, r.shade_grid, s)) != hipSuccess ||
        (err = wave_fill(p, r.shade_grid, s)) != hipSuccess || (err = mark(0)) != hipSuccess)
        return err;
    r.launches += 5;

    hipGraphExec_t exec = nullptr;
    if (!eager) {
        std::vector<uint8_t> key(sizeof(WaveParams) + 4 * sizeof(int));
        std::memcpy(key.data(), &p, sizeof(WaveParams));
        int extra[4] = {r.trace_grid, r.shade_grid, K, r.count ? 1 : 0};
        std::memcpy(key.data() + sizeof(WaveParams), extra, sizeof(extra));
        for (auto& g : e.graphs)
            if (g.key == key) exec = g.exec;
        if (!exec) {
            if (e.graphs.size() >= 8) wave_drop_graphs(e);
            if ((err = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess) return err;
            hipError_t le = iterations();
            hipGraph_t graph = nullptr;
            err = hipStreamEndCapture(s, &graph);
            if (le != hipSuccess) err = le;
            if (err == hipSuccess) err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (err != hipSuccess) return err;
            WaveEngine::Graph g;
            g.key = std::move(key);
            g.exec = exec;
            e.graphs.push_back(std::move(g));
        }
    }
    // every path ends within max_depth bounces plus the iteration that
    // finishes it, and every iteration refills the slots freed by the one
    // before, so this many replays always suffice
    const int64_t max_replays =
        ((r.total_samples + p.pool - 1) / p.pool + 2) * (int64_t)(p.max_depth + 2) / K + 2;
    auto replay = [&](int slot) -> hipError_t {
        hipError_t x = exec ? hipGraphLaunch(exec, s) : iterations();
        if (x != hipSuccess) return x;
        r.launches += K * per_iter;
        r.iterations += K;
        int* dst = e.host_counts + slot * kPollInts;
        if ((x = hipMemcpyAsync(dst, p.ctr + (size_t)kCtrQueue * kCtrStride, kPollInts * sizeof(int),
                                hipMemcpyDeviceToHost, s)) != hipSuccess)
            return x;
        return hipEventRecord(e.poll[slot], s);
    };
    // after an even number of iterations the next one is at parity 0: it
    // traces queue 0 and finishes the paths listed at parity 1
    auto live = [&](int slot) {
        const int* src = e.host_counts + slot * kPollInts;
        int64_t n = 0;
        for (int q = 0; q < kQueueRegions * kShards; q++) n += src[(size_t)(queue_ctr(0, 0, 0) + q) * kCtrStride];
        for (int k = 0; k < kShards; k++) n += src[(size_t)(kCtrTerm + kShards + k) * kCtrStride];
        return n;
    };
    if ((err = replay(0)) != hipSuccess) return err;
    // The sky shading is read only by the accumulation: it runs on the side
    // stream from the end of the first replay on, beside the window's later
    // iterations -- the tail kernel's few waves leave most of the ALUs idle.
    if ((err = hipEventRecord(e.side_start, s)) != hipSuccess ||
        (err = hipStreamWaitEvent(e.side, e.side_start, 0)) != hipSuccess ||
        (err = wave_sky_shade(p, r.sky_grid, e.side)) != hipSuccess ||
        (err = hipEventRecord(e.side_done, e.side)) != hipSuccess)
        return err;
    for (int64_t i = 0;; i++) {
        if (i + 1 >= max_replays) return hipErrorLaunchTimeOut;  // cannot happen for a valid window
        if ((err = replay((int)((i + 1) & 1))) != hipSuccess) return err;
        if ((err = hipEventSynchronize(e.poll[i & 1])) != hipSuccess) return err;
        if (live((int)(i & 1)) == 0) break;
    }
    if ((err = hipStreamWaitEvent(s, e.side_done, 0)) != hipSuccess ||
        (err = wave_accumulate(p, r.npix_total, s)) != hipSuccess || (err = mark(4)) != hipSuccess)
        return err;
    r.launches += 1;
    return collect();
}

}  // namespace oxr
