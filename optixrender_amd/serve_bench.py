"""Serving benchmark: render latency under continuous batching, plain
accumulation against the denoiser at equal image error.

Workload (seeded):
  * resident scenes: the Cornell box (2,572 tris), the sphere field (983,040),
    the heightfield with the medium box (20,016,980) and a 5 M-triangle
    sphere field;
  * Poisson arrivals at `rate` requests/s; resolution uniform over 256^2,
    512^2, 1024^2 and 1920x1080; spp log-uniform in [1, 64]; depth 16.

Two arms over the same request stream:
  * plain: every request at its spp;
  * denoised: every request at spp / k with the KPCNN, where k is the
    equal-RMSE factor -- plain accumulation needs k times the samples of a
    denoised render to reach the same RMSE of the tone-mapped image against
    a converged reference (measured by `equal_rmse_factor` on the sphere
    field at 1080p, by interpolation between measured plain sample counts).

Latency is the server's own submit-to-done time per request (p50/p99).

    python -m optixrender_amd.serve_bench --requests 400 --rate 200
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


def run_arm(ctx: Context, resident, stream, denoised: bool, k: float, budget_gb: float = 160.0,
            tick_paths: int = 1 << 26, depth: int = 16):
    srv = Server(ctx, budget_bytes=int(budget_gb * (1 << 30)), max_active=64, tick_paths=tick_paths,
                 max_frame_pixels=1920 * 1080)
    ids = []
    paths = 0
    try:
        t0 = time.perf_counter()
        for r in stream:
            delay = t0 + r["t"] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            name, sid, cam = resident[r["scene"]]
            w, h = r["res"]
            spp = max(1, int(round(r["spp"] / k))) if denoised else r["spp"]
            ids.append(srv.submit(sid, w, h, spp, cam, max_depth=depth, rr_start=0, denoise=denoised))
            paths += w * h * spp
        for rid in ids:
            if srv.wait(rid, 600000) != 0:
                raise RuntimeError(f"request {rid} did not finish")
        wall = time.perf_counter() - t0
        st = srv.stats()
        return {"requests": len(ids), "p50_ms": round(st.p50_ms, 2), "p99_ms": round(st.p99_ms, 2),
                "mean_ms": round(st.mean_ms, 2), "wall_s": round(wall, 2), "ticks": int(st.ticks),
                "Mpaths_s": round(paths / wall / 1e6, 1), "paths": int(paths)}
    finally:
        for rid in ids:
            try:
                srv.release(rid)
            except abi.OxrError:
                pass
        srv.close()


def display(rgb: np.ndarray) -> np.ndarray:
    """The tone map of ComputeFinalPixelColors (PostProcessing.cu:26-48) in
    float, before the 8-bit quantisation: what a viewer of the image sees."""
    c = np.asarray(rgb, np.float64)[..., :3]
    lum = 0.299 * c[..., 0] + 0.587 * c[..., 1] + 0.114 * c[..., 2]
    f = 1.0 / (1.0 + lum / 1.5)
    return np.clip(np.maximum(c * f[..., None], 0.0) ** (1.0 / 2.2), 0.0, 1.0)


def display_rmse(a, b) -> float:
    return float(np.sqrt(np.mean((display(a) - display(b)) ** 2)))


def equal_rmse_factor(ctx: Context, sid: int, cam, w=1920, h=1080, spp_dn=4, ref_spp=4096, depth=16,
                      plain_spp=(4, 8, 16, 32, 64, 128, 256)):
    """Samples plain accumulation needs, relative to spp_dn denoised samples,
    for the same RMSE of the displayed (tone-mapped) image against a ref_spp
    render: log-log interpolation between the measured plain sample counts."""
    def render(spp, dn=False):
        ctx.select(sid)
        ctx.resize(w, h)
        ctx.set_camera(cam)
        ctx.render(spp=spp, max_depth=depth, rr_start=0, denoise=dn)
        return ctx.download_aov(3 if dn else 0, w, h)

    ref = render(ref_spp)
    plain = {s: display_rmse(render(s), ref) for s in plain_spp}
    dn = display_rmse(render(spp_dn, True), ref)
    xs = np.log(np.array(plain_spp, float))
    ys = np.log(np.array([plain[s] for s in plain_spp]))
    ly = math.log(dn)
    if ly >= ys[0]:
        spp_equal = float(plain_spp[0]) * math.exp((ys[0] - ly) / (ys[1] - ys[0]) * (xs[1] - xs[0]))
    elif ly <= ys[-1]:
        spp_equal = float("inf")  # better than the largest plain count measured
    else:
        i = int(np.searchsorted(-ys, -ly))  # ys decreasing
        f = (ly - ys[i - 1]) / (ys[i] - ys[i - 1])
        spp_equal = float(math.exp(xs[i - 1] + f * (xs[i] - xs[i - 1])))
    factor = spp_equal / spp_dn if math.isfinite(spp_equal) else float(plain_spp[-1]) / spp_dn
    return {"metric": "RMSE of the tone-mapped image", "ref_spp": ref_spp, "denoised_spp": spp_dn,
            "rmse_denoised": round(dn, 5), "rmse_plain": {str(k): round(v, 5) for k, v in plain.items()},
            "plain_spp_at_equal_rmse": round(spp_equal, 1) if math.isfinite(spp_equal) else f">{plain_spp[-1]}",
            "factor": round(factor, 2)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--small", action="store_true", help="only the Cornell box and the 1 M sphere field resident")
    ap.add_argument("--factor", type=float, default=0.0, help="equal-RMSE factor (0: measure it)")
    args = ap.parse_args(argv)
    ctx = Context(0)
    ctx.load_denoiser(denoise.flat_weights(denoise.load_weights()))
    resident = upload_scenes(ctx, full=not args.small)
    eq = None
    k = args.factor
    if k <= 0:
        _, sid, cam = resident[1]
        eq = equal_rmse_factor(ctx, sid, cam)
        k = eq["factor"]
        log(f"equal RMSE: {json.dumps(eq)}")
    stream = request_stream(args.requests, args.rate, len(resident))
    plain = run_arm(ctx, resident, stream, False, 1.0)
    log(f"plain: {json.dumps(plain)}")
    dn = run_arm(ctx, resident, stream, True, k)
    log(f"denoised: {json.dumps(dn)}")
    out = {"requests": args.requests, "rate_per_s": args.rate, "scenes": [r[0] for r in resident],
           "equal_rmse": eq, "factor": k, "plain": plain, "denoised": dn}
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
