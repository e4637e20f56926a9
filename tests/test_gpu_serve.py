"""Render server (continuous batching) on the GPU: every request's image is
the one a single render of that request gives, whatever it shared the device
with; admission respects the HBM budget; impossible requests are refused."""
import numpy as np
import pytest

from optixrender_amd import abi, denoise, scenes
from optixrender_amd.renderer import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cornell(ctx):
    sc = scenes.cornell_spheres()
    sid = ctx.upload(sc)
    ctx.build_bvh(sid)
    yield sc, sid
    ctx.release(sid)


def single(ctx, sid, w, h, spp, cam, depth=8):
    ctx.select(sid)
    ctx.resize(w, h)
    ctx.set_camera(cam)
    ctx.render(spp=spp, max_depth=depth, rr_start=0)
    return ctx.download_aov(0, w, h), ctx.download_rgba8(w, h)


REQS = [
    (40, 32, 3, ((0.0, 1.0, 3.0), (0.0, 1.0, 0.0))),
    (64, 48, 5, ((0.3, 1.2, 2.8), (0.0, 0.8, 0.0))),
    (24, 24, 1, ((-0.4, 0.9, 2.5), (0.2, 1.0, -0.5))),
    (32, 40, 7, ((0.0, 1.5, 3.0), (0.0, 0.5, 0.0))),
]


def test_each_result_equals_its_single_render(ctx, cornell):
    sc, sid = cornell
    srv = Server(ctx, budget_bytes=1 << 30, max_active=8, tick_paths=3000, max_frame_pixels=64 * 64)
    try:
        ids = [srv.submit(sid, w, h, spp, abi.camera(*cam), max_depth=8) for w, h, spp, cam in REQS]
        for rid in ids:
            assert srv.wait(rid, 60000) == 0
        got = [(srv.result_color(rid), srv.result_rgba8(rid)) for rid in ids]
        st = srv.stats()
        assert st.completed == len(REQS)
        assert st.ticks >= 2              # work was split over several shared ticks
        assert st.p99_ms >= st.p50_ms > 0
    finally:
        srv.close()
    for (w, h, spp, cam), (col, px) in zip(REQS, got):
        ref_col, ref_px = single(ctx, sid, w, h, spp, abi.camera(*cam))
        assert np.array_equal(col.view(np.uint32), ref_col.view(np.uint32))
        assert np.array_equal(px, ref_px)


def test_budget_admission_serialises(ctx, cornell):
    sc, sid = cornell
    tick, fp = 2048, 64 * 64
    probe = Server(ctx, budget_bytes=1 << 34, max_active=8, tick_paths=tick, max_frame_pixels=fp)
    wave = probe.stats().bytes_in_use          # idle server: the shared wavefront buffers only
    probe.close()
    one = 64 * 48 * 56 + 8 * 6                   # colour + albedo + normal float4, depth float, RGBA8; 8x8 cover tiles
    srv = Server(ctx, budget_bytes=wave + one + 100, max_active=8, tick_paths=tick, max_frame_pixels=fp)
    try:
        ids = [srv.submit(sid, 64, 48, 2, abi.camera((0, 1, 3), (0, 1, 0)), max_depth=8) for _ in range(3)]
        # a finished request keeps its buffers until released, so each one is
        # released once waited for; the next one is admitted only then
        for rid in ids:
            st = srv.stats()
            assert st.bytes_in_use <= st.budget_bytes and st.active <= 1
            assert srv.wait(rid, 60000) == 0
            assert srv.stats().active == 0
            srv.release(rid)
        assert srv.stats().completed == 3
        with pytest.raises(abi.OxrError):
            srv.submit(sid, 200, 200, 1, abi.camera((0, 1, 3), (0, 1, 0)))   # larger than max_frame_pixels
        assert srv.stats().rejected == 1
    finally:
        srv.close()


def test_cancel(ctx, cornell):
    sc, sid = cornell
    srv = Server(ctx, budget_bytes=1 << 30, max_active=1, tick_paths=1024, max_frame_pixels=64 * 64)
    try:
        a = srv.submit(sid, 64, 64, 64, abi.camera((0, 1, 3), (0, 1, 0)), max_depth=8)
        b = srv.submit(sid, 32, 32, 2, abi.camera((0, 1, 3), (0, 1, 0)), max_depth=8)
        srv.cancel(a)
        assert srv.wait(b, 60000) == 0
        assert srv.poll(a) == abi.REQ_CANCELLED
        assert srv.stats().cancelled == 1
    finally:
        srv.close()


def test_denoised_request_equals_denoised_render(ctx, cornell):
    sc, sid = cornell
    ctx.load_denoiser(denoise.flat_weights(denoise.load_weights()))
    w, h, spp, cam = REQS[1]
    srv = Server(ctx, budget_bytes=1 << 30, max_active=4, tick_paths=2000, max_frame_pixels=64 * 64)
    try:
        a = srv.submit(sid, w, h, spp, abi.camera(*cam), max_depth=8, denoise=True)
        b = srv.submit(sid, *REQS[0][:3], abi.camera(*REQS[0][3]), max_depth=8)   # shares the ticks, plain
        assert srv.wait(a, 60000) == 0 and srv.wait(b, 60000) == 0
        col, px = srv.result_color(a), srv.result_rgba8(a)
    finally:
        srv.close()
    ctx.select(sid)
    ctx.resize(w, h)
    ctx.set_camera(abi.camera(*cam))
    ctx.render(spp=spp, max_depth=8, rr_start=0, denoise=True)
    assert np.array_equal(col.view(np.uint32), ctx.download_aov(3, w, h).view(np.uint32))
    assert np.array_equal(px, ctx.download_rgba8(w, h))



def test_ticks_of_several_scenes_overlap_and_equal_single_renders(ctx, cornell):
    # requests over different resident scenes in one server: a textured
    # scene, a lit sphere field (sky pass, lights in view) and the Cornell
    # box.  A tick serves one group (scene, depth, RR start, camera medium);
    # the runners' ticks of different groups run at the same time, each on
    # its own stream.  Each result equals its single render.
    _, cid = cornell
    tex = scenes.textured_cornell(subdiv=2)
    lit = scenes.lit_sphere_field(subdiv=1)
    tid, lid = ctx.upload(tex), ctx.upload(lit)
    ctx.build_bvh(tid)
    ctx.build_bvh(lid)
    reqs = [(tid, 40, 32, 24, tex.camera), (lid, 48, 27, 32, lit.camera),
            (cid, 32, 32, 32, scenes.cornell_spheres().camera), (tid, 24, 24, 40, tex.camera)]
    srv = Server(ctx, budget_bytes=1 << 31, max_active=8, tick_paths=2000, max_frame_pixels=64 * 64, runners=3)
    try:
        ids = [srv.submit(s, w, h, spp, cam, max_depth=8) for s, w, h, spp, cam in reqs]
        for rid in ids:
            assert srv.wait(rid, 60000) == 0
        got = [srv.result_color(rid) for rid in ids]
        st = srv.stats()
        print(f"[serve] {st.ticks} ticks on {st.runners} runners, {st.overlapped_ticks} began beside another, "
              f"at most {st.max_running} at once", flush=True)
        assert st.runners == 3
        assert st.max_running >= 2 and st.overlapped_ticks >= 1
    finally:
        srv.close()
    try:
        for (s, w, h, spp, cam), col in zip(reqs, got):
            ref, _ = single(ctx, s, w, h, spp, cam)
            assert np.array_equal(col.view(np.uint32), ref.view(np.uint32))
    finally:
        ctx.release(tid)
        ctx.release(lid)


def test_cancel_and_release_during_a_tick(ctx, cornell):
    # A request cancelled and released while its tick is running (and one
    # still pending) is only forgotten by the caller's side; the worker frees
    # it when it drops it.  The server keeps serving afterwards, and the
    # scene cannot be released while a request of it is live.
    sc, sid = cornell
    cam = abi.camera((0, 1, 3), (0, 1, 0))
    srv = Server(ctx, budget_bytes=1 << 33, max_active=1, tick_paths=1 << 22, max_frame_pixels=256 * 256)
    try:
        a = srv.submit(sid, 256, 256, 64, cam, max_depth=8)     # 4 M paths: one long tick
        b = srv.submit(sid, 64, 64, 8, cam, max_depth=8)        # pending behind a (max_active 1)
        with pytest.raises(abi.OxrError):
            ctx.release(sid)                                     # live requests hold the scene
        for _ in range(2000):
            if srv.poll(a) == abi.REQ_ACTIVE:
                break
        srv.cancel(a)
        srv.cancel(b)
        srv.release(a)
        srv.release(b)
        with pytest.raises(abi.OxrError):
            srv.poll(a)                                          # forgotten at once
        with pytest.raises(abi.OxrError):
            srv.poll(b)
        c = srv.submit(sid, 40, 32, 3, cam, max_depth=8)
        assert srv.wait(c, 60000) == 0
        col = srv.result_color(c)
        st = srv.stats()
        assert st.cancelled == 2 and st.completed == 1
        srv.release(c)
    finally:
        srv.close()
    ref, _ = single(ctx, sid, 40, 32, 3, cam)
    assert np.array_equal(col.view(np.uint32), ref.view(np.uint32))


def test_scene_release_waits_for_server_requests(ctx):
    sc = scenes.cornell_spheres(subdiv=1)
    sid = ctx.upload(sc)
    ctx.build_bvh(sid)
    srv = Server(ctx, budget_bytes=1 << 30, max_active=2, tick_paths=4096, max_frame_pixels=64 * 64)
    try:
        rid = srv.submit(sid, 48, 48, 4, sc.camera, max_depth=8)
        assert srv.wait(rid, 60000) == 0
        # once done, the worker has let go of the request: the scene can go,
        # and the finished image stays readable
        ctx.release(sid)
        assert srv.result_rgba8(rid).shape == (48, 48)
    finally:
        srv.close()
