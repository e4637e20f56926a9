"""oxr_trace_rays_noalloc (device rays and results, scratch in a fixed device
array, no allocation) against oxr_trace_rays on the same rays: hits and
occlusion bit for bit, two calls in a row with the second one smaller, so
that what the first call left in the shared scratch cannot leak into the
second.  Device buffers here are the test's own (oxr_malloc)."""
import ctypes as C

import numpy as np
import pytest

from optixrender_amd import abi, scenes

pytestmark = pytest.mark.gpu


def make_rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 1e-4
    r[:, 4:7] = d
    r[:, 7] = 3.4e38
    r[::5, 7] = 3.0                 # some finite tmax
    return r


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.p = C.c_void_p()
        ctx.check(ctx._lib.oxr_malloc(ctx.h, max(nbytes, 4), C.byref(self.p)), "oxr_malloc")

    def upload(self, a):
        a = np.ascontiguousarray(a)
        self.ctx.check(self.ctx._lib.oxr_memcpy_h2d(self.ctx.h, self.p, abi.ptr(a), a.nbytes), "h2d")

    def download(self, a):
        self.ctx.check(self.ctx._lib.oxr_memcpy_d2h(self.ctx.h, abi.ptr(a), self.p, a.nbytes), "d2h")
        return a

    def free(self):
        self.ctx._lib.oxr_free(self.ctx.h, self.p)


def test_noalloc_trace_equals_trace(ctx):
    sc = scenes.sphere_field(subdiv=3, env_w=64, env_h=32)
    lo, hi = sc.positions.reshape(-1, 3).min(0), sc.positions.reshape(-1, 3).max(0)
    sizes = [300_000, 4_099]        # more rays than the fixed grid's lanes, then a smaller call
    sid = ctx.upload(sc)
    bufs = []
    try:
        ctx.build_bvh(sid)
        ctx.select(sid)
        d_rays, d_hits, d_occ = DeviceBuffer(ctx, 32 * sizes[0]), DeviceBuffer(ctx, 16 * sizes[0]), \
            DeviceBuffer(ctx, sizes[0])
        bufs = [d_rays, d_hits, d_occ]
        for k, n in enumerate(sizes):
            print(f"[noalloc] call {k + 1}: {n} rays", flush=True)
            rays = make_rays(n, 100 + k, lo, hi)
            d_rays.upload(rays)
            rc = ctx._lib.oxr_trace_rays_noalloc(ctx.h, d_rays.p, n, 0, d_hits.p, None)
            ctx.check(rc, "oxr_trace_rays_noalloc")
            got = d_hits.download(np.zeros((n, 4), np.float32))
            rc = ctx._lib.oxr_trace_rays_noalloc(ctx.h, d_rays.p, n, 1, None, d_occ.p)
            ctx.check(rc, "oxr_trace_rays_noalloc")
            got_occ = d_occ.download(np.zeros(n, np.uint8))
            ref, _ = ctx.trace(rays)
            ref_occ, _ = ctx.trace(rays, any_hit=True)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
            assert np.array_equal(got_occ.astype(bool), ref_occ)
            assert (got[:, 1].view(np.uint32) != 0xFFFFFFFF).mean() > 0.05    # the rays do hit things
        # sizes beyond the maximum are refused before any device work
        rc = ctx._lib.oxr_trace_rays_noalloc(ctx.h, d_rays.p, (1 << 31) + 1, 0, d_hits.p, None)
        assert rc == 1                  # OXR_ERR_INVALID
        assert ctx._lib.oxr_trace_rays_noalloc(ctx.h, d_rays.p, 0, 0, d_hits.p, None) == 0
    finally:
        for b in bufs:
            b.free()
        ctx.release(sid)
