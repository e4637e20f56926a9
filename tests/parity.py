"""Shared checks of the GPU parity tests: a GPU render and an oracle render
of the same scene, compared path by path.

Per-path bar: radiance within 1e-4 relative (plus 1e-7 absolute) on at least
99.9% of paths; a path that is NaN on one side must be NaN on the other.  The
rest are hit points an ulp apart (device vs host transcendentals in BSDF
sampling) whose paths then diverge.
"""
import numpy as np

import oracle_lib
from optixrender_amd import abi

REL_TOL = 1e-4
PASS_FRACTION = 0.999


def per_path_check(gpu, ref, label):
    a = gpu.reshape(-1, 3).astype(np.float64)
    b = ref.reshape(-1, 3).astype(np.float64)
    nan_a = np.isnan(a).any(1)
    nan_b = np.isnan(b).any(1)
    ok = np.all(np.abs(a - b) <= REL_TOL * np.maximum(np.abs(a), np.abs(b)) + 1e-7, axis=1)
    ok = np.where(nan_a | nan_b, nan_a == nan_b, ok)
    frac = ok.mean()
    print(f"[{label}] paths={ok.size} within 1e-4: {frac:.6f}  diverged={int((~ok).sum())}", flush=True)
    return frac


def _render_pair(ctx, scene, w, h, spp, depth, rr, pool=None, tail=None, normals=None, filt=abi.FILTER_GAUSSIAN):
    """GPU render and oracle render of the same scene.  pool: path slots of
    the GPU pool (None = default); a pool smaller than spp*w*h restarts
    every slot many times.  tail: the tail-mode threshold (None = default;
    0 keeps every path in the per-material kernels).  normals: a list that
    receives the (GPU, oracle) normal AOVs.  filt: the pixel filter."""
    sid = ctx.upload(scene)
    base = ctx.get_tuning()
    try:
        ctx.build_bvh(sid)
        ctx.select(sid)
        ctx.resize(w, h)
        ctx.set_camera(scene.camera)
        if pool is not None or tail is not None:
            t = ctx.get_tuning()
            if pool is not None:
                t.pool_paths = pool
            if tail is not None:
                t.tail_paths = tail
            ctx.set_tuning(t)
        ctx.render(spp=spp, max_depth=depth, rr_start=rr, filter=filt, flags=abi.OXR_FLAG_KEEP_PATH_RADIANCE)
        paths = ctx.download_path_radiance(spp, w, h)
        color = ctx.download_aov(0, w, h)
        albedo = ctx.download_aov(1, w, h)
        normal = ctx.download_aov(2, w, h)
        dz = ctx.download_depth(w, h)
    finally:
        ctx.set_tuning(base)
        ctx.release(sid)
    rc, ra, rn, rd, rp = oracle_lib.render(scene, w, h, spp, max_depth=depth, rr_start=rr, filt=filt,
                                           keep_paths=True, with_depth=True)
    # first-hit depth AOV: same traversal, same distance arithmetic
    close = np.abs(dz - rd) <= 1e-6 * np.abs(rd) + 1e-7
    assert close.mean() >= PASS_FRACTION, f"depth AOV differs on {np.count_nonzero(~close)} pixels"
    if normals is not None:
        normals.extend([normal, rn])
    return paths, rp, color, rc, albedo, ra
