"""GPU vs oracle at the benchmark workloads themselves, through the C ABI.

The smaller parity tests use reduced scenes; these run the configurations at
their own geometry and camera:

  * Cornell box with two icospheres (2,572 triangles), 256 x 256, 16 spp,
    depth 8, Russian roulette from bounce 3: the oracle's full config.
  * Sphere field, 12 icospheres at subdivision 6 (983,040 triangles) under
    the 2048 x 1024 HDR sky: LBVH bit-exact (58% of its triangles share a
    30-bit Morton code, so the index tie-break builds most of the lower
    tree), and 1920 x 1080 x 2 spp at depth 16 per path.
  * fBm heightfield with the medium box (20,016,980 triangles): LBVH
    bit-exact, and 480 x 270 x 2 spp at depth 32 per path.
  * The serving mix: requests drawn from the seeded Poisson stream over the
    four resident scenes (the three above and the 5 M-triangle sphere
    field) at its resolutions and spp, each equal to its single render.

Per-path bar as everywhere: within 1e-4 relative on at least 99.9% of paths.
"""
import numpy as np
import pytest

import oracle_lib
from parity import PASS_FRACTION, _render_pair, per_path_check
from optixrender_amd import abi, scenes, serve_bench
from optixrender_amd.renderer import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sphere_field():
    return scenes.sphere_field()


@pytest.fixture(scope="module")
def heightfield():
    return scenes.heightfield_medium()


def _lbvh_equal(ctx, sc, label, rounds=-1):
    n = sc.n_tris
    print(f"[lbvh] {label} n={n} rounds={rounds}", flush=True)
    sid = ctx.upload(sc)
    try:
        ctx.build_bvh(sid, rounds)
        codes, order, left, right, lo, hi = ctx.download_bvh(sid, n)
        qnode = ctx.download_nodes(sid, n)
    finally:
        ctx.release(sid)
    ref = oracle_lib.lbvh(sc.positions, rounds)
    assert np.array_equal(codes, ref["codes"])
    assert np.array_equal(order, ref["order"])
    assert np.array_equal(left, ref["left"])
    assert np.array_equal(right, ref["right"])
    assert np.array_equal(lo.view(np.uint32), ref["lo"].view(np.uint32))
    assert np.array_equal(hi.view(np.uint32), ref["hi"].view(np.uint32))
    assert np.array_equal(qnode, ref["qnode"])
    shared = 1.0 - np.unique(codes).size / n
    print(f"[lbvh] {label}: bit-exact; {shared:.1%} of triangles share a Morton code", flush=True)


@pytest.mark.parametrize("rounds", [0, -1])
def test_lbvh_bit_exact_sphere_field(ctx, sphere_field, rounds):
    _lbvh_equal(ctx, sphere_field, "sphere field 983,040 tris", rounds)


def test_lbvh_bit_exact_heightfield(ctx, heightfield):
    _lbvh_equal(ctx, heightfield, "heightfield 20,016,980 tris")


def test_cornell_config_per_path(ctx):
    sc = scenes.cornell_spheres()
    print("[workload] cornell 256x256x16 depth 8 rr 3", flush=True)
    paths, rp, *_ = _render_pair(ctx, sc, 256, 256, 16, 8, 3)
    assert per_path_check(paths, rp, "cornell 256^2 x16") >= PASS_FRACTION


def test_sphere_field_1080p_per_path(ctx, sphere_field):
    print("[workload] sphere field 1920x1080x2 depth 16", flush=True)
    paths, rp, *_ = _render_pair(ctx, sphere_field, 1920, 1080, 2, 16, 0)
    assert per_path_check(paths, rp, "sphere field 1080p x2") >= PASS_FRACTION


def test_sphere_field_1080p_ray_counts(ctx, sphere_field):
    # The traversal roofline counts node fetches and triangle tests in a
    # counting render of the GPU; the oracle counts the same quantities with
    # its canonical full-stack traversal.  Paths that diverge (under 0.1%)
    # shift the totals a little, nothing more.
    w, h, spp = 1920, 1080, 2
    sid = ctx.upload(sphere_field)
    try:
        ctx.build_bvh(sid)
        ctx.select(sid)
        ctx.resize(w, h)
        ctx.set_camera(sphere_field.camera)
        st = ctx.render(spp=spp, max_depth=16, rr_start=0, flags=abi.OXR_FLAG_COUNT_TRAVERSAL)
    finally:
        ctx.release(sid)
    ref = oracle_lib.render_stats(sphere_field, w, h, spp, max_depth=16, rr_start=0)
    gpu = dict(extension_rays=st.extension_rays, shadow_rays=st.shadow_rays, node_fetches=st.node_fetches,
               tri_tests=st.tri_tests, ext_node_fetches=st.ext_node_fetches, ext_tri_tests=st.ext_tri_tests)
    print(f"[counts] gpu {gpu}\n[counts] oracle {ref}", flush=True)
    for k, v in gpu.items():
        assert abs(v - ref[k]) <= 1e-4 * ref[k], k
    assert st.extension_rays >= w * h * spp      # every camera ray is traced when counting


def test_heightfield_per_path(ctx, heightfield):
    print("[workload] heightfield 20M tris 480x270x2 depth 32", flush=True)
    paths, rp, *_ = _render_pair(ctx, heightfield, 480, 270, 2, 32, 0)
    assert per_path_check(paths, rp, "heightfield 480x270 x2") >= PASS_FRACTION


def test_serving_mix_equals_single_renders(ctx, sphere_field, heightfield):
    # 20 requests of the seeded serving stream (Poisson arrivals, resolutions
    # over 256^2 .. 1920x1080, spp log-uniform in [1, 64], depth 16) over the
    # four resident scenes, submitted at once so that ticks are shared and
    # the ticks of different scenes run at the same time on four runners.
    resident = [scenes.cornell_spheres(), sphere_field, heightfield, scenes.sphere_field_5m()]
    sids = []
    try:
        for sc in resident:
            sid = ctx.upload(sc)
            ctx.build_bvh(sid)
            sids.append(sid)
        stream = serve_bench.request_stream(20, 200.0, len(resident))
        srv = Server(ctx, budget_bytes=64 << 30, max_active=32, tick_paths=1 << 25, max_frame_pixels=1920 * 1080,
                     runners=4)
        try:
            ids = []
            for r in stream:
                w, h = r["res"]
                ids.append(srv.submit(sids[r["scene"]], w, h, r["spp"], resident[r["scene"]].camera, max_depth=16))
            for rid in ids:
                assert srv.wait(rid, 300000) == 0
            got = [srv.result_rgba8(rid) for rid in ids]
            cols = [srv.result_color(rid) for rid in ids]
            st = srv.stats()
            print(f"[serve] {len(ids)} requests, {st.ticks} ticks ({st.overlapped_ticks} beside another, at most "
                  f"{st.max_running} at once), {st.paths / 1e6:.1f} M paths, "
                  f"p50 {st.p50_ms:.1f} ms p99 {st.p99_ms:.1f} ms", flush=True)
            assert st.max_running >= 2
        finally:
            srv.close()
        for r, px, col in zip(stream, got, cols):
            w, h = r["res"]
            ctx.select(sids[r["scene"]])
            ctx.resize(w, h)
            ctx.set_camera(resident[r["scene"]].camera)
            ctx.render(spp=r["spp"], max_depth=16, rr_start=0)
            assert np.array_equal(col.view(np.uint32), ctx.download_aov(0, w, h).view(np.uint32)), r
            assert np.array_equal(px, ctx.download_rgba8(w, h)), r
    finally:
        for sid in sids:
            ctx.release(sid)
