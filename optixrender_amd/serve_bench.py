"""Serving benchmark: render latency under continuous batching, plain
accumulation against the denoiser at equal image error.

Workload (seeded):
  * resident scenes: the Cornell box (2,572 tris), the sphere field (983,040),
    the heightfield with the medium box (20,016,980) and a 5 M-triangle
    sphere field;
  * Poisson arrivals at `rate` requests/s; resolution uniform over 256^2,
    512^2, 1024^2 and 1920x1080; spp log-uniform in [1, 64]; depth 16.

Equal error, per scene.  For each resident scene the display RMSE (the tone
map of ComputeFinalPixelColors in float, before 8-bit quantisation) against
a 4096-spp plain render is measured at 512^2 for plain accumulation at 1 ..
256 spp and for the denoiser at 1 .. 64 spp.  A request asking for s plain
samples is served in the denoised arm by the fewest denoised samples s' < s
whose RMSE is at most the plain RMSE at s (log-log interpolation between the
measured plain counts); when no such s' exists -- the network does not beat
plain accumulation on that scene at that sample count -- it is served plain,
the denoiser's fallback.

Two arms over the same request stream: plain (every request at its spp) and
denoised (the plan above); then plain again over a quarter as many requests
at three quarters of the rate the plain arm actually served, a load it
sustains, so that its latency is render time rather than queueing.  Latency is the server's own submit-to-done time
per request; the time spent waiting for admission is reported beside it.
Each arm's offered load is compared with the scenes' single-render
throughput: an arm whose offered paths per second exceed what the GPU renders
is overloaded, and its latency is then mostly queueing.

    python -m optixrender_amd.serve_bench --requests 10000 --rate 200 --summary out.md
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import abi, denoise, scenes
from .renderer import Context, Server

RESOLUTIONS = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
CURVE_RES = (512, 512)
REF_SPP = 4096
PLAIN_SPP = (1, 2, 4, 8, 16, 32, 64, 128, 256)
DN_SPP = (1, 2, 4, 8, 16, 32, 64)
DEPTH = 16


def log(msg):
    print(f"[serve] {msg}", file=sys.stderr, flush=True)


def upload_scenes(ctx: Context, full: bool = True):
    """(name, scene id, camera) of the resident scenes."""
    gens = [("cornell", lambda: scenes.cornell_spheres()),
            ("sphere_field", lambda: scenes.sphere_field(seed=1, subdiv=6))]
    if full:
        gens += [("heightfield_medium", lambda: scenes.heightfield_medium()),
                 ("sphere_field_5m", lambda: scenes.sphere_field_5m())]
    out = []
    for name, g in gens:
        t0 = time.perf_counter()
        sc = g()
        sid = ctx.upload(sc)
        ms = ctx.build_bvh(sid)
        log(f"scene {name}: {sc.n_tris} tris, generated+uploaded in {time.perf_counter() - t0:.1f}s, BVH {ms:.1f} ms")
        out.append((name, sid, sc.camera))
        del sc
    return out


def request_stream(n: int, rate: float, n_scenes: int, seed: int = 1):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    res = rng.integers(0, len(RESOLUTIONS), n)
    spp = np.clip(np.round(np.exp(rng.uniform(0.0, math.log(64.0), n))), 1, 64).astype(int)
    scene = rng.integers(0, n_scenes, n)
    return [dict(t=float(t[i]), res=RESOLUTIONS[res[i]], spp=int(spp[i]), scene=int(scene[i])) for i in range(n)]


def display(rgb: np.ndarray) -> np.ndarray:
    """The tone map of ComputeFinalPixelColors (PostProcessing.cu:26-48) in
    float, before the 8-bit quantisation: what a viewer of the image sees."""
    c = np.asarray(rgb, np.float64)[..., :3]
    lum = 0.299 * c[..., 0] + 0.587 * c[..., 1] + 0.114 * c[..., 2]
    f = 1.0 / (1.0 + lum / 1.5)
    return np.clip(np.maximum(c * f[..., None], 0.0) ** (1.0 / 2.2), 0.0, 1.0)


def display_rmse(a, b) -> float:
    return float(np.sqrt(np.mean((display(a) - display(b)) ** 2)))


def scene_curves(ctx: Context, sid: int, cam, w: int = CURVE_RES[0], h: int = CURVE_RES[1], ref_spp: int = REF_SPP,
                 depth: int = DEPTH):
    """Display RMSE against a ref_spp plain render: plain at PLAIN_SPP,
    denoised at DN_SPP (same frames 0..spp-1 as a served request)."""
    def render(spp, dn=False):
        ctx.select(sid)
        ctx.resize(w, h)
        ctx.set_camera(cam)
        ctx.render(spp=spp, max_depth=depth, rr_start=0, denoise=dn)
        return ctx.download_aov(3 if dn else 0, w, h)

    ref = render(ref_spp)
    plain = {s: display_rmse(render(s), ref) for s in PLAIN_SPP}
    dn = {s: display_rmse(render(s, True), ref) for s in DN_SPP}
    return {"resolution": [w, h], "ref_spp": ref_spp, "plain": plain, "denoised": dn}


def plain_rmse_at(curve: dict, s: float) -> float:
    """Plain RMSE at s samples: log-log interpolation between measured counts
    (extrapolated with the last segment's slope beyond them)."""
    xs = np.log(np.array(sorted(curve["plain"]), float))
    ys = np.log(np.array([curve["plain"][k] for k in sorted(curve["plain"])]))
    x = math.log(s)
    i = int(np.clip(np.searchsorted(xs, x) - 1, 0, len(xs) - 2))
    f = (x - xs[i]) / (xs[i + 1] - xs[i])
    return float(math.exp(ys[i] + f * (ys[i + 1] - ys[i])))


def equal_rmse_plan(curve: dict, max_spp: int = 64):
    """spp -> (denoise, spp served, plain RMSE target, RMSE served) for plain
    requests of 1..max_spp samples."""
    plan = {}
    for s in range(1, max_spp + 1):
        target = plain_rmse_at(curve, s)
        ok = [k for k in sorted(curve["denoised"]) if k < s and curve["denoised"][k] <= target]
        if ok:
            plan[s] = (True, ok[0], target, curve["denoised"][ok[0]])
        else:
            plan[s] = (False, s, target, target)
    return plan


def single_render_rates(ctx: Context, resident, w: int = 1920, h: int = 1080, spp: int = 16, depth: int = DEPTH):
    """Mpaths/s of one plain render of each resident scene (what the GPU
    renders of that scene with nothing else on it)."""
    out = {}
    for name, sid, cam in resident:
        ctx.select(sid)
        ctx.resize(w, h)
        ctx.set_camera(cam)
        ctx.render(spp=2, max_depth=depth, rr_start=0)            # warm
        st = ctx.render(spp=spp, max_depth=depth, rr_start=0)
        out[name] = round(w * h * spp / (st.render_ms / 1e3) / 1e6, 1)
    return out


def offered_load(resident, stream, plans, rates, rate):
    """Offered paths/s per arm and the utilisation it implies: the sum over
    requests of paths / (the scene's single-render rate), per second of
    arrivals."""
    span = stream[-1]["t"] - stream[0]["t"] if len(stream) > 1 else 1.0
    out = {}
    for arm in ("plain", "denoised"):
        paths = busy = 0.0
        for r in stream:
            name = resident[r["scene"]][0]
            w, h = r["res"]
            spp = plans[name][r["spp"]][1] if arm == "denoised" else r["spp"]
            p = w * h * spp
            paths += p
            busy += p / (rates[name] * 1e6)
        out[arm] = {"offered_Mpaths_s": round(paths / span / 1e6, 1), "utilisation": round(busy / span, 3)}
    return out


def percentile(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(q * (len(v) - 1) + 0.5))]


def run_arm(ctx: Context, resident, stream, plan_of, budget_gb: float = 160.0, tick_paths: int = 1 << 26,
            depth: int = DEPTH, runners: int = 2):
    """Submit the stream at its arrival times; plan_of(request) -> (denoise,
    spp).  A reaper thread releases every finished request as soon as it is
    done (a finished request holds its buffers, and so its share of the HBM
    budget, until released), after reading the server's per-request times
    (submit -> done, and the part spent waiting for admission)."""
    import threading

    srv = Server(ctx, budget_bytes=int(budget_gb * (1 << 30)), max_active=64, tick_paths=tick_paths,
                 max_frame_pixels=1920 * 1080, runners=runners)
    meta = {}            # id -> (scene name, denoised)
    times = {}           # id -> (queued ms, total ms)
    lock = threading.Lock()
    outstanding = []
    submitted_all = threading.Event()
    failed = []

    def reap():
        while True:
            with lock:
                ids = list(outstanding)
            if not ids and submitted_all.is_set():
                return
            for rid in ids:
                st = srv.poll(rid)
                if st == abi.REQ_DONE:
                    times[rid] = srv.times(rid)
                    srv.release(rid)
                    with lock:
                        outstanding.remove(rid)
                elif st in (abi.REQ_FAILED, abi.REQ_CANCELLED):
                    failed.append(rid)
                    srv.release(rid)
                    with lock:
                        outstanding.remove(rid)
            time.sleep(0.002)

    reaper = threading.Thread(target=reap, daemon=True)
    paths = 0
    try:
        reaper.start()
        t0 = time.perf_counter()
        for i, r in enumerate(stream):
            delay = t0 + r["t"] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            name, sid, cam = resident[r["scene"]]
            w, h = r["res"]
            dn, spp = plan_of(r)
            rid = srv.submit(sid, w, h, spp, cam, max_depth=depth, rr_start=0, denoise=dn)
            meta[rid] = (name, dn)
            with lock:
                outstanding.append(rid)
            paths += w * h * spp
            if (i + 1) % 1000 == 0:
                log(f"  submitted {i + 1}/{len(stream)}, {len(times)} done")
        submitted_all.set()
        last = time.perf_counter()
        while reaper.is_alive():   # progress while the backlog drains
            reaper.join(15.0)
            if reaper.is_alive():
                log(f"  draining: {len(times)}/{len(stream)} done at {time.perf_counter() - t0:.0f}s")
            if time.perf_counter() - last > 1200:
                raise RuntimeError("serving arm did not drain")
        wall = time.perf_counter() - t0
        st = srv.stats()
    finally:
        submitted_all.set()
        srv.close()
    if failed:
        raise RuntimeError(f"{len(failed)} requests failed")
    total = [t for _, t in times.values()]
    queued = [q for q, _ in times.values()]
    per_scene = {}
    for rid, (q, t) in times.items():
        per_scene.setdefault(meta[rid][0], []).append(t)
    return {"requests": len(times), "p50_ms": round(percentile(total, 0.5), 2), "p99_ms": round(percentile(total, 0.99), 2),
            "mean_ms": round(float(np.mean(total)), 2), "queued_p50_ms": round(percentile(queued, 0.5), 2),
            "queued_p99_ms": round(percentile(queued, 0.99), 2), "wall_s": round(wall, 2), "ticks": int(st.ticks),
            "runners": int(st.runners), "overlapped_ticks": int(st.overlapped_ticks),
            "max_ticks_at_once": int(st.max_running),
            "Mpaths_s": round(paths / wall / 1e6, 1), "paths": int(paths),
            "denoised_requests": int(sum(1 for _, dn in meta.values() if dn)),
            "per_scene_p50_p99_ms": {k: [round(percentile(v, 0.5), 1), round(percentile(v, 0.99), 1)]
                                     for k, v in per_scene.items()}}


def measure(ctx: Context, resident, requests: int, rate: float, seed: int = 1):
    """Curves, plans, single-render rates and both arms."""
    curves, plans = {}, {}
    for name, sid, cam in resident:
        t0 = time.perf_counter()
        curves[name] = scene_curves(ctx, sid, cam)
        plans[name] = equal_rmse_plan(curves[name])
        log(f"curves {name} ({time.perf_counter() - t0:.1f}s): plain "
            f"{ {k: round(v, 4) for k, v in curves[name]['plain'].items()} } denoised "
            f"{ {k: round(v, 4) for k, v in curves[name]['denoised'].items()} }")
    rates = single_render_rates(ctx, resident)
    log(f"single-render Mpaths/s at 1080p x16: {rates}")
    stream = request_stream(requests, rate, len(resident), seed)
    load = offered_load(resident, stream, plans, rates, rate)
    log(f"offered load: {load}")
    plain = run_arm(ctx, resident, stream, lambda r: (False, r["spp"]))
    log(f"plain: {json.dumps(plain)}")
    dn = run_arm(ctx, resident, stream, lambda r: plans[resident[r["scene"]][0]][r["spp"]][:2])
    log(f"denoised: {json.dumps(dn)}")
    # Plain accumulation at a load it sustains: three quarters of the rate
    # the overloaded plain arm actually served, over a quarter as many
    # requests of the same stream law.  Its latency is render time rather
    # than queueing.
    sus_rate = 0.75 * rate * plain["Mpaths_s"] / max(load["plain"]["offered_Mpaths_s"], 1e-9)
    sus_stream = request_stream(max(requests // 4, 1), sus_rate, len(resident), seed + 1)
    sus = run_arm(ctx, resident, sus_stream, lambda r: (False, r["spp"]))
    sus["rate_per_s"] = round(sus_rate, 2)
    log(f"plain at a sustained load: {json.dumps(sus)}")
    served = {name: {s: {"denoise": p[0], "spp": p[1]} for s, p in plan.items() if s in (1, 2, 4, 8, 16, 32, 64)}
              for name, plan in plans.items()}
    return {"requests": requests, "rate_per_s": rate, "scenes": [r[0] for r in resident],
            "curves": {k: {"resolution": v["resolution"], "ref_spp": v["ref_spp"],
                           "plain": {str(s): round(x, 5) for s, x in v["plain"].items()},
                           "denoised": {str(s): round(x, 5) for s, x in v["denoised"].items()}}
                       for k, v in curves.items()},
            "plan": served, "single_render_Mpaths_s": rates, "offered": load, "plain": plain, "denoised": dn,
            "plain_sustained": sus,
            "plain_overloaded": load["plain"]["utilisation"] >= 1.0,
            "denoised_overloaded": load["denoised"]["utilisation"] >= 1.0}


def summary_markdown(res: dict, command: str) -> str:
    lines = [f"# Serving: {res['requests']} Poisson requests at {res['rate_per_s']:.0f}/s over four resident scenes",
             "", "Command (repository root, one MI355X):", "", "```", command, "```", "",
             "Equal error per scene: display RMSE against a 4096-spp plain render at 512^2 (PLAIN at 1..256 spp, "
             "DENOISED at 1..64 spp); a plain request of s samples is served denoised at the fewest s' < s samples "
             "whose RMSE is at most plain's at s, else plain.", "",
             "| scene | single-render Mpaths/s | plain RMSE 4 / 16 / 64 spp | denoised RMSE 1 / 4 / 16 spp | "
             "plan for 8 / 32 / 64 spp |", "|---|---:|---|---|---|"]
    for name in res["scenes"]:
        c = res["curves"][name]
        pl = " / ".join(f"{c['plain'][k]:.4f}" for k in ("4", "16", "64"))
        dn = " / ".join(f"{c['denoised'][k]:.4f}" for k in ("1", "4", "16"))
        plan = " / ".join(("dn %d" % res["plan"][name][s]["spp"]) if res["plan"][name][s]["denoise"]
                          else ("plain %d" % s) for s in (8, 32, 64))
        lines.append(f"| {name} | {res['single_render_Mpaths_s'][name]} | {pl} | {dn} | {plan} |")
    lines += ["", "| arm | offered Mpaths/s | utilisation | served Mpaths/s | p50 ms | p99 ms | queued p50 / p99 ms | "
              "denoised requests |", "|---|---:|---:|---:|---:|---:|---|---:|"]
    for arm in ("plain", "denoised"):
        a, o = res[arm], res["offered"][arm]
        lines.append(f"| {arm} | {o['offered_Mpaths_s']} | {o['utilisation']} | {a['Mpaths_s']} | {a['p50_ms']} | "
                     f"{a['p99_ms']} | {a['queued_p50_ms']} / {a['queued_p99_ms']} | {a['denoised_requests']} |")
    a = res["plain_sustained"]
    lines.append(f"| plain, {a['requests']} requests at {a['rate_per_s']}/s | - | - | {a['Mpaths_s']} | {a['p50_ms']} | "
                 f"{a['p99_ms']} | {a['queued_p50_ms']} / {a['queued_p99_ms']} | 0 |")
    lines += ["", f"Runners (ticks in flight at once, each with its own path pool and stream): {res['plain']['runners']}; "
              f"ticks that began beside another: plain {res['plain']['overlapped_ticks']} of {res['plain']['ticks']}, "
              f"denoised {res['denoised']['overlapped_ticks']} of {res['denoised']['ticks']}."]
    lines += ["", "Per scene p50 / p99 ms:", "", "| scene | plain | denoised |", "|---|---|---|"]
    for name in res["scenes"]:
        p, d = res["plain"]["per_scene_p50_p99_ms"][name], res["denoised"]["per_scene_p50_p99_ms"][name]
        lines.append(f"| {name} | {p[0]} / {p[1]} | {d[0]} / {d[1]} |")
    lines += ["", "Full result (JSON):", "", "```", json.dumps(res), "```", ""]
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=10000)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--summary", default="", help="write a markdown summary here")
    args = ap.parse_args(argv)
    ctx = Context(0)
    ctx.load_denoiser(denoise.flat_weights(denoise.load_weights()))
    resident = upload_scenes(ctx)
    res = measure(ctx, resident, args.requests, args.rate)
    print(json.dumps(res), flush=True)
    if args.summary:
        cmd = f"python3 -m optixrender_amd.serve_bench --requests {args.requests} --rate {args.rate:g}"
        with open(args.summary, "w") as f:
            f.write(summary_markdown(res, cmd))
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
