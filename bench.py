"""Benchmark of the path-tracing hot path on one MI355X.

One step = one render call of the headline workload: the 983,040-triangle
sphere field (12 icospheres at subdivision 6 under a 2048x1024 procedural HDR
sky) at 1920x1080, 64 spp, max depth 16: 132.7 M paths through the
regenerating wavefront loop (trace -> regenerate -> shade on a resident path
pool, then accumulate and tone map).  Scene upload and the BVH build happen
before the timed region (the build time is reported separately).

Prints one JSON line with Mpaths/s, the traversal roofline of the trace
kernel (algorithmic bytes from an untimed counting pass of the same frames,
kernel time from HIP events on the library's stream during the timed steps,
HBM bytes from the committed counter summary) and the CPU oracle timed on
this host's cores.  Legs beside the headline (in `config`): the viewer loop,
the denoiser, serving, and the batched configuration -- the 20 M-triangle
heightfield with the medium box at 1080p, 32 spp, depth 32, autotuned and
replayed from captured graphs, with its own roofline and CPU oracle sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES, RAY_BYTES = 16, 36, 48   # 32-B record per two child boxes, 36-B triangle, ray in + hit out
SURVEY_NODE_BYTES = 32   # one uncompressed 32-B box record per child (lo, ref, hi), the per-ray formula's unit
SHADOW_RAY_BYTES = 40   # 4-B slot id + 32-B ray in + 4-B flag cleared when occluded


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=16)
    ap.add_argument("--subdiv", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-heightfield", action="store_true",
                    help="skip the 20 M-triangle heightfield leg (1080p, 32 spp, depth 32, autotuned graphs)")
    ap.add_argument("--workload", choices=["sphere_field", "heightfield"], default="sphere_field",
                    help="with --profile-steps: which workload's steps to run")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg (p50/p99 latency at equal RMSE)")
    ap.add_argument("--no-viewer", action="store_true", help="skip the viewer leg (1-spp frames with readback)")
    ap.add_argument("--no-denoiser", action="store_true", help="skip the denoiser leg (KPCNN time per 1080p frame)")
    ap.add_argument("--profile-denoiser", action="store_true",
                    help="only the denoiser leg: 4-spp renders denoised at the bench resolution (for rocprofv3)")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests per serving arm (10,000 in optixrender_amd.serve_bench's own run)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--profile-steps", action="store_true", help="only run the timed steps (for rocprofv3)")
    ap.add_argument("--tuning", default="",
                    help="comma-separated oxr_tuning fields to set before the steps, e.g. stack_lds=16,pool_paths=4194304")
    ap.add_argument("--autotune", action="store_true", help="run oxr_autotune on an 8-spp probe before the steps")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device copy (the known-size reference for HBM counters)")
    return ap.parse_args()


def host_cpu():
    """The host's CPU as the OS reports it: logical CPUs, the ones this
    process may run on, and the model name."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)), "model": model}


def oracle_rate(scene, W, H, depth, threads, spp, row_step=1):
    """Mpaths/s of the CPU oracle over W x H (every row_step-th row) at spp
    frames: paths traced over the time spent tracing them, after the oracle's
    own LBVH build and environment cache (reported separately)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib  # checker only: the baseline leg, never the measured path

    t0 = time.perf_counter()
    st = oracle_lib.render_stats(scene, W, H, spp, max_depth=depth, threads=threads, row_step=row_step)
    wall = time.perf_counter() - t0
    return st["paths"] / st["trace_s"] / 1e6, st, wall - st["trace_s"]


def cpu_baseline(scene, W, H, depth, spp_total, label, budget_s=12.0, single_s=4.0):
    """The CPU oracle on this host: every usable core, and one core, each on
    a bounded sample of the same workload sized from a quick probe (whole
    frames when they fit the budget, else every k-th row of one frame; the
    camera and per-pixel RNG are those of the full workload)."""
    cores = len(os.sched_getaffinity(0))
    probe_rate, _, setup_s = oracle_rate(scene, W, H, depth, cores, 1, row_step=16)
    full_frame_s = W * H / (probe_rate * 1e6)
    if full_frame_s <= budget_s:
        spp, step = int(max(1, min(spp_total, budget_s // full_frame_s))), 1
    else:
        spp, step = 1, int(math.ceil(full_frame_s / budget_s))
    rate, st, _ = oracle_rate(scene, W, H, depth, cores, spp, row_step=step)
    one_step = int(max(1, math.ceil(W * H / (probe_rate / cores * 1e6) / single_s)))
    rate1, st1, _ = oracle_rate(scene, W, H, depth, 1, 1, row_step=one_step)
    frac = st["paths"] / (W * H * spp_total)
    return {
        "value": round(rate, 4),
        "unit": "Mpaths/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{label}: {W}x{H}" + (f", every {step}th row" if step > 1 else "") +
                  f", {spp} spp (frames 0..{spp - 1}) of the {spp_total}-spp workload: {st['paths']} paths, "
                  f"{frac:.4f} of the workload (time extrapolates linearly by {1 / frac:.1f}x); "
                  f"{st['extension_rays']} extension rays",
        "seconds": round(st["trace_s"], 2),
        "single_thread_Mpaths_s": round(rate1, 4),
        "single_thread_sample": f"{W}x{H}" + (f", every {one_step}th row" if one_step > 1 else "") +
                                f", 1 spp: {st1['paths']} paths in {st1['trace_s']:.2f} s",
        "setup_s_excluded": round(setup_s, 2),
        "host": host_cpu(),
    }


NODE_BYTES_DOC = ("16 B per child box fetched (a 32-B compressed record holds both children of a node), 36-B "
                  "triangle per test, 48 B per extension ray, 40 per shadow ray")


def traversal_bytes(counted, node_bytes=NODE_BYTES):
    """Algorithmic bytes of one step's k_trace launches from a counting
    render of the same frames: camera rays the cover mask rules out are never
    traced, and the tail kernel's rays are its own launch."""
    NODE_BYTES = node_bytes
    ext_bytes = NODE_BYTES * (counted.ext_node_fetches - counted.resolved_node_fetches) + \
        TRI_BYTES * (counted.ext_tri_tests - counted.resolved_tri_tests) + \
        RAY_BYTES * (counted.extension_rays - counted.camera_rays_resolved)
    sh_nodes = counted.node_fetches - counted.ext_node_fetches
    sh_tris = counted.tri_tests - counted.ext_tri_tests
    sh_bytes = NODE_BYTES * sh_nodes + TRI_BYTES * sh_tris + SHADOW_RAY_BYTES * counted.shadow_rays
    tail_bytes = NODE_BYTES * counted.tail_node_fetches + TRI_BYTES * counted.tail_tri_tests + \
        RAY_BYTES * counted.tail_extension_rays + SHADOW_RAY_BYTES * counted.tail_shadow_rays
    ext_rays = max(counted.extension_rays - counted.camera_rays_resolved, 1)
    return ext_bytes + sh_bytes - tail_bytes, ext_bytes / ext_rays, sh_bytes / max(counted.shadow_rays, 1)


def roofline(counted, trace_ms, launches, traffic_file, workload, tuning, strict=True):
    """The roofline object of k_trace: `achieved` / `frac` from algorithmic
    bytes (every node and triangle fetch at full size), `traffic` the HBM
    bytes per launch the counters measured (FETCH_SIZE + WRITE_SIZE, scaled
    by a copy of known size), and `dram_frac` those bytes over the same
    kernel time: how much of the algorithmic traffic reaches DRAM."""
    alg, per_ext, per_sh = traversal_bytes(counted)
    achieved = alg / (trace_ms / 1e3) / 1e9
    measured = measured_traffic(traffic_file, workload, tuning, strict)
    out = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured["per_launch"] if measured else None,
        "kernel": "k_trace (all launches of one step: extension and shadow rays)",
        "launches_per_step": int(launches), "bytes_per_step": int(alg),
        "bytes_basis": f"algorithmic: {NODE_BYTES_DOC}, node and triangle counts from a counting render",
        "per_ray_bytes": {"extension": round(per_ext, 1), "shadow": round(per_sh, 1)},
        # active lanes per lane slot of the traversal loop's iterations
        "simd_use": round(counted.trace_lane_visits / max(counted.trace_wave_slots, 1), 4),
        "simd_use_basis": "node visits of active lanes / (64 x the most visits of any lane of each 64-ray chunk), "
                          "from the counting render, which also traces the camera rays the cover mask resolves",
    }
    # the same counts priced at 32 B per child box, the unit of uncompressed
    # (lo, ref, hi) records: how fast the traversal covers the tree
    alg32 = traversal_bytes(counted, SURVEY_NODE_BYTES)[0]
    out["uncompressed_equivalent"] = {"achieved": round(alg32 / (trace_ms / 1e3) / 1e9, 1),
                                      "frac": round(alg32 / (trace_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                                      "basis": "32 B per child box, 36 B per triangle, 48 / 40 B per ray"}
    if measured is not None:
        step_bytes = measured["per_step"] or measured["per_launch"] * launches
        dram = step_bytes / (trace_ms / 1e3) / 1e9
        out["dram_achieved"] = round(dram, 1)
        out["dram_frac"] = round(dram / HBM_PEAK_GBS, 4)
        out["traffic_source"] = f"{traffic_file.relative_to(ROOT)} (tuning {measured['tuning']})"
    return out


def serving(ctx, args):
    """The latency half of the metric: Poisson arrivals over four resident
    scenes (Cornell box, this sphere field, the 20 M-triangle heightfield
    with the medium box, a 5 M-triangle sphere field) served by the
    continuous-batching server, plain accumulation against the KPCNN
    denoiser at equal display RMSE per scene (optixrender_amd.serve_bench:
    each request's denoised sample count comes from its scene's measured
    RMSE curves, plain accumulation where the network does not win)."""
    from optixrender_amd import denoise, serve_bench

    print("[bench] serving: uploading the resident scenes", file=sys.stderr, flush=True)
    ctx.load_denoiser(denoise.flat_weights(denoise.load_weights()))
    resident = serve_bench.upload_scenes(ctx)
    try:
        res = serve_bench.measure(ctx, resident, args.serve_requests, args.serve_rate)
    finally:
        for _, s, _ in resident:
            ctx.release(s)
    res.update({"plain_p50_ms": res["plain"]["p50_ms"], "plain_p99_ms": res["plain"]["p99_ms"],
                "denoised_p50_ms": res["denoised"]["p50_ms"], "denoised_p99_ms": res["denoised"]["p99_ms"],
                "plain_sustained_p50_ms": res["plain_sustained"]["p50_ms"],
                "plain_sustained_p99_ms": res["plain_sustained"]["p99_ms"]})
    return res


COPY_BYTES = 1 << 30


def copy_calibration():
    """A plain device-to-device copy of COPY_BYTES (reads and writes exactly
    that many bytes): under rocprofv3 --pmc its FETCH_SIZE / WRITE_SIZE give
    the factor that turns those counters into HBM bytes on this GPU."""
    import torch

    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[bench] copy calibration: 3 copies of {COPY_BYTES} bytes", file=sys.stderr, flush=True)
    del src, dst


TRAFFIC_FILE = ROOT / "profiles" / "pmc_k_trace_sphere_field_1080p_64spp.md"
TRAFFIC_FILE_HF = ROOT / "profiles" / "pmc_k_trace_heightfield_1080p_32spp.md"


def tuning_key(t) -> str:
    return json.dumps(t.as_dict(), sort_keys=True)


def measured_traffic(path: Path, workload: str, tuning: str, strict: bool = True):
    """HBM bytes of k_trace from a committed counter summary (lines
    `workload: ...`, `tuning: ...`, `traffic_bytes_per_launch: N` and
    optionally `traffic_bytes_per_step: N`) measured on this workload: a dict
    (per_launch, per_step, tuning), or None.  strict: only when its tuning is
    this run's too (per-launch bytes depend on the launch shape)."""
    if not path.exists():
        return None
    fields = {}
    for line in path.read_text().splitlines():
        if ":" in line and not line.startswith(("|", "#", " ", "`")):
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    if fields.get("workload") != workload or (strict and fields.get("tuning") != tuning):
        return None
    if "traffic_bytes_per_launch" not in fields:
        return None
    per_step = fields.get("traffic_bytes_per_step")
    return {"per_launch": int(float(fields["traffic_bytes_per_launch"])),
            "per_step": int(float(per_step)) if per_step else None, "tuning": fields.get("tuning")}


def viewer(ctx, args, scene, frames=32):
    """The interactive viewer's loop (main.cpp: Render() then DownloadPixels()
    every frame): progressive 1-spp frames at the bench resolution, read back
    by the synchronous copy into pageable memory, and by the streamed,
    double-buffered page-locked readback (OXR_FLAG_PRESENT)."""
    from optixrender_amd import abi

    W, H = args.width, args.height
    out = {"frames": frames, "spp_per_frame": 1}
    ctx.set_camera(scene.camera)
    ctx.render(spp=1, max_depth=args.depth, progressive=True, flags=abi.OXR_FLAG_PRESENT)   # warm: graphs, buffers
    ctx.present(copy=False)
    for name, flags in (("download", 0), ("present", abi.OXR_FLAG_PRESENT)):
        ctx.set_camera(scene.camera)
        render_ms = 0.0
        t0 = time.perf_counter()
        for _ in range(frames):
            st = ctx.render(spp=1, max_depth=args.depth, progressive=True, flags=flags)
            render_ms += st.render_ms
            px = ctx.download_rgba8(W, H) if name == "download" else ctx.present(copy=False)[0]
        wall = (time.perf_counter() - t0) * 1e3 / frames
        out[f"{name}_ms_per_frame"] = round(wall, 3)
        out[f"{name}_device_ms_per_frame"] = round(render_ms / frames, 3)
        del px
    out["fps_download"] = round(1e3 / out["download_ms_per_frame"], 1)
    out["fps_present"] = round(1e3 / out["present_ms_per_frame"], 1)
    print(f"[bench] viewer: {json.dumps(out)}", file=sys.stderr, flush=True)
    return out


DN_FLOP_PER_PIXEL = 75_456 + 150   # five 3x3 convolutions (2*9*Cin*32 each) + the 5x5 kernel application
MFMA_BF16_PEAK_TFLOPS = 2500.0     # MI355X dense BF16


def denoiser(ctx, args, scene, frames=8):
    """The KPCNN on the bench frame: 4-spp renders with denoise = 1; the
    render statistics time the five BF16 MFMA convolutions and the kernel
    application with HIP events (denoise_ms)."""
    from optixrender_amd import denoise

    W, H = args.width, args.height
    ctx.load_denoiser(denoise.flat_weights(denoise.load_weights()))
    ctx.set_camera(scene.camera)
    ms = []
    for i in range(frames + 1):
        st = ctx.render(spp=4, max_depth=args.depth, denoise=True)
        if i > 0:
            ms.append(st.denoise_ms)
    avg = float(np.mean(ms))
    tflops = DN_FLOP_PER_PIXEL * W * H / (avg / 1e3) / 1e12
    out = {"frames": frames, "pixels": W * H, "ms_per_frame": round(avg, 4), "flop_per_pixel": DN_FLOP_PER_PIXEL,
           "achieved_tflops": round(tflops, 1), "peak_tflops_bf16": MFMA_BF16_PEAK_TFLOPS,
           "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4)}
    print(f"[bench] denoiser: {json.dumps(out)}", file=sys.stderr, flush=True)
    return out


HF_W, HF_H, HF_SPP, HF_DEPTH = 1920, 1080, 32, 32
# autotune probes: 8 spp of 1080p (16.6 M samples) fill the largest pool
PROBE_SPP = HF_PROBE_SPP = 8


def heightfield_leg(ctx, args, steps=3):
    """The batched configuration: 1920x1080, 32 spp, depth 32 over the
    20,016,980-triangle fBm heightfield with the homogeneous-medium box and
    the dielectric sphere, per-triangle interleaved materials.  The launch
    tuning comes from oxr_autotune on an 8-spp probe; the timed steps replay
    the captured iteration graphs.  Its own counting render gives the
    traversal roofline, and the CPU oracle is timed on a sample of it."""
    from optixrender_amd import abi, scenes

    t0 = time.perf_counter()
    sc = scenes.heightfield_medium()
    gen_s = time.perf_counter() - t0
    sid = ctx.upload(sc)
    build_ms = ctx.build_bvh(sid)
    base = ctx.get_tuning()
    try:
        ctx.select(sid)
        ctx.resize(HF_W, HF_H)
        ctx.set_camera(sc.camera)
        print(f"[bench] heightfield: {sc.n_tris} tris generated in {gen_s:.1f}s, BVH build {build_ms:.1f} ms",
              file=sys.stderr, flush=True)
        # an 8-spp probe: 16.6 M samples fill the largest pool, so that pool
        # sizes are compared on a full pool as in the 32-spp step
        tuned, probe_ms = ctx.autotune(spp=HF_PROBE_SPP, max_depth=HF_DEPTH, rr_start=0)
        ctx.set_camera(sc.camera)
        print(f"[bench] heightfield: autotune {tuned.as_dict()} ({probe_ms:.2f} ms per {HF_PROBE_SPP}-spp probe)",
              file=sys.stderr, flush=True)

        def step(flags):
            return ctx.render(spp=HF_SPP, max_depth=HF_DEPTH, rr_start=0, progressive=False, flags=flags)

        counted = step(abi.OXR_FLAG_COUNT_TRAVERSAL)
        staged = [step(abi.OXR_FLAG_STAGE_TIMING) for _ in range(2)][-1]
        step(0)                                     # warm the graphs
        tg = time.perf_counter()
        for _ in range(steps):
            step(0)
        ms = (time.perf_counter() - tg) * 1e3 / steps
        paths = HF_W * HF_H * HF_SPP
        workload = f"heightfield_medium {sc.n_tris} tris, {HF_W}x{HF_H}, {HF_SPP} spp, depth {HF_DEPTH} (seed 1)"
        out = {
            "workload": workload, "metric": "Mpaths/s", "value": round(paths / (ms / 1e3) / 1e6, 3),
            "ms_per_step": round(ms, 3), "steps": steps, "launch": "captured iteration graphs, autotuned",
            "bvh_build_ms": round(build_ms, 3), "tuning": tuned.as_dict(), "autotune_probe_ms": round(probe_ms, 3),
            "stage_ms": {"trace": round(staged.trace_ms, 3), "shade": round(staged.shade_ms, 3),
                         "regenerate": round(staged.generate_ms, 3), "accumulate": round(staged.accumulate_ms, 3),
                         "eager_total": round(staged.render_ms, 3)},
            "iterations_per_step": int(staged.iterations), "extension_rays": int(counted.extension_rays),
            "shadow_rays": int(counted.shadow_rays),
            "roofline": roofline(counted, staged.trace_ms, staged.iterations, TRAFFIC_FILE_HF, workload,
                                 tuning_key(ctx.get_tuning()), strict=False),
        }
        print(f"[bench] heightfield: {ms:.1f} ms/step, {out['value']} Mpaths/s", file=sys.stderr, flush=True)
        if not args.no_cpu_baseline:
            print("[bench] heightfield: timing the CPU oracle sample", file=sys.stderr, flush=True)
            out["cpu_baseline"] = cpu_baseline(sc, HF_W, HF_H, HF_DEPTH, HF_SPP, "heightfield_medium")
        return out
    finally:
        ctx.set_tuning(base)
        ctx.release(sid)


def profile_heightfield(args):
    """Only the heightfield leg's graph-replayed steps (for rocprofv3): the
    tuning is --tuning's, else the default one."""
    from optixrender_amd import scenes
    from optixrender_amd.renderer import Context

    sc = scenes.heightfield_medium()
    ctx = Context(0)
    sid = ctx.upload(sc)
    ctx.build_bvh(sid)
    ctx.select(sid)
    ctx.resize(HF_W, HF_H)
    ctx.set_camera(sc.camera)
    if args.tuning:
        t = ctx.get_tuning()
        for kv in args.tuning.split(","):
            k, v = kv.split("=")
            setattr(t, k.strip(), int(v))
        ctx.set_tuning(t)
    print(f"[bench] heightfield profile: tuning {tuning_key(ctx.get_tuning())}", file=sys.stderr, flush=True)
    for i in range(args.warmup + args.steps):
        st = ctx.render(spp=HF_SPP, max_depth=HF_DEPTH, rr_start=0, progressive=False, flags=0)
        print(f"[bench] heightfield profile step {i + 1}: {st.render_ms:.1f} ms, {st.iterations} iterations",
              file=sys.stderr, flush=True)
    ctx.close()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the renderer runs on one GPU only; --gpus must be 1 (sharding a frame "
                                   "over several GPUs by image tiles is designed, not built)"}))
        sys.exit(2)
    from optixrender_amd import abi, scenes
    from optixrender_amd.renderer import Context

    if args.copy_calibration:
        copy_calibration()

    if args.profile_steps and args.workload == "heightfield":
        profile_heightfield(args)
        return

    t0 = time.perf_counter()
    scene = scenes.sphere_field(seed=1, subdiv=args.subdiv)
    gen_s = time.perf_counter() - t0
    ctx = Context(0)
    sid = ctx.upload(scene)
    build_ms = ctx.build_bvh(sid)
    ctx.select(sid)
    ctx.resize(args.width, args.height)
    ctx.set_camera(scene.camera)
    W, H, spp, depth = args.width, args.height, args.spp, args.depth
    print(f"[bench] scene {scene.n_tris} tris generated in {gen_s:.1f}s, BVH build {build_ms:.2f} ms", file=sys.stderr,
          flush=True)
    if args.tuning:
        t = ctx.get_tuning()
        for kv in args.tuning.split(","):
            k, v = kv.split("=")
            setattr(t, k.strip(), int(v))
        ctx.set_tuning(t)
    if args.autotune:
        t, ms = ctx.autotune(spp=PROBE_SPP, max_depth=depth, rr_start=0)
        ctx.set_camera(scene.camera)
        print(f"[bench] autotune: {t.as_dict()} ({ms:.2f} ms per {PROBE_SPP}-spp probe)", file=sys.stderr, flush=True)

    if args.profile_denoiser:
        print(json.dumps({"denoiser": denoiser(ctx, args, scene)}), flush=True)
        ctx.close()
        return

    def step(flags=abi.OXR_FLAG_STAGE_TIMING):
        # stage timing: kernels launched one by one with HIP events between stages
        return ctx.render(spp=spp, max_depth=depth, rr_start=0, progressive=False, flags=flags)

    counted = None
    if not args.profile_steps:
        counted = step(abi.OXR_FLAG_COUNT_TRAVERSAL)   # untimed: algorithmic bytes of the same frames
        print(f"[bench] counting pass done: {counted.extension_rays} ext rays, {counted.shadow_rays} shadow rays",
              file=sys.stderr, flush=True)
    # The timed steps are the product path: each render replays its captured
    # iteration graphs.  (--profile-steps: eager launches, one kernel at a
    # time, for the profiler.)
    timed_flags = abi.OXR_FLAG_STAGE_TIMING if args.profile_steps else 0
    for i in range(args.warmup):
        step(timed_flags)
        print(f"[bench] warmup {i + 1}/{args.warmup}", file=sys.stderr, flush=True)
    ctx_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = step(timed_flags)
        ctx_ms.append(st.render_ms)
        print(f"[bench] step {i + 1}/{args.steps}: {st.render_ms:.1f} ms ({st.iterations} iterations)",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    # Stage times from eager launches with HIP events between the stages
    # (the roofline's kernel time): a few more steps, untimed by the wall clock.
    trace_ms, shade_ms, regen_ms, eager_ms = [], [], [], []
    for i in range(0 if args.profile_steps else 3):
        sst = step()
        eager_ms.append(sst.render_ms)
        trace_ms.append(sst.trace_ms)
        shade_ms.append(sst.shade_ms)
        regen_ms.append(sst.generate_ms)
        print(f"[bench] staged step {i + 1}: {sst.render_ms:.1f} ms (trace {sst.trace_ms:.1f}, shade "
              f"{sst.shade_ms:.1f}, regenerate {sst.generate_ms:.1f}, accumulate "
              f"{sst.accumulate_ms:.1f}; {sst.iterations} iterations, {sst.kernel_launches} launches)",
              file=sys.stderr, flush=True)
    if not trace_ms:
        trace_ms, shade_ms, regen_ms, eager_ms = [st.trace_ms], [st.shade_ms], [st.generate_ms], [st.render_ms]
    paths = W * H * spp
    value = paths / (wall / args.steps) / 1e6
    out = {
        "metric": "Mpaths/s 1080p 1M-tri; p50/p99 render latency at equal RMSE vs plain",
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"sphere_field {scene.n_tris} tris, {W}x{H}, {spp} spp, depth {depth}, "
                               f"2048x1024 HDR sky (seed 1)",
                   "bvh_build_ms": round(build_ms, 3), "device_ms_per_step": round(float(np.mean(ctx_ms)), 3),
                   "trace_ms_per_step": round(float(np.mean(trace_ms)), 3),
                   "shade_ms_per_step": round(float(np.mean(shade_ms)), 3),
                   "regenerate_ms_per_step": round(float(np.mean(regen_ms)), 3),
                   "iterations_per_step": int(st.iterations), "launches_per_step": int(st.kernel_launches),
                   "tuning": ctx.get_tuning().as_dict(),
                   "launch": ("eager launches with HIP events between stages" if args.profile_steps else
                              "captured iteration graphs (the product path); stage times from eager steps with "
                              "HIP events between stages"),
                   "eager_ms_per_step": round(float(np.mean(eager_ms)), 3)},
    }
    if counted is not None:
        out["roofline"] = roofline(counted, float(np.mean(trace_ms)), st.iterations, TRAFFIC_FILE,
                                   out["config"]["workload"], tuning_key(ctx.get_tuning()))
        out["config"].update({"extension_rays": int(counted.extension_rays), "shadow_rays": int(counted.shadow_rays),
                              "camera_rays_resolved_without_trace": int(counted.camera_rays_resolved),
                              "tail_kernel_rays": int(counted.tail_extension_rays + counted.tail_shadow_rays),
                              "traced_rays_per_s": round((counted.extension_rays - counted.camera_rays_resolved +
                                                          counted.shadow_rays) /
                                                  (ms_per_step / 1e3), 0)})
    if not args.no_viewer and not args.profile_steps:
        out["config"]["viewer"] = viewer(ctx, args, scene)
    if not args.no_denoiser and not args.profile_steps:
        out["config"]["denoiser"] = denoiser(ctx, args, scene)
    if not args.no_serving and not args.profile_steps:
        out["config"]["serving"] = serving(ctx, args)
    if not args.no_heightfield and not args.profile_steps:
        out["config"]["heightfield"] = heightfield_leg(ctx, args)
    if not args.no_cpu_baseline and not args.profile_steps:
        print("[bench] timing the CPU oracle sample", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(scene, W, H, depth, spp, "sphere_field")
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
