"""Camera-ray culling: a sample whose pixel tile lies outside the image of
every box of the BVH's top levels is shaded by the window's sky pass without
a trace (its camera ray can meet only the environment or an analytic light).
The render must not change by a single bit.

Cases: the sphere field from its bench camera (most of the image is sky),
the same scene seen from among the spheres (boxes straddle the camera plane)
and from a camera looking away from it (every tile ruled out), the Cornell box
(the camera is inside the geometry: nothing can be ruled out), and the
heightfield with the medium box.  Every pixel filter, since the tile margin
covers the largest filter offset."""
import numpy as np
import pytest

from optixrender_amd import abi, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def field(ctx):
    sc = scenes.sphere_field(seed=1, subdiv=3, env_w=256, env_h=128)
    sid = ctx.upload(sc)
    ctx.build_bvh(sid)
    yield sc, sid
    ctx.release(sid)


@pytest.fixture(scope="module")
def cornell(ctx):
    sc = scenes.cornell_spheres()
    sid = ctx.upload(sc)
    ctx.build_bvh(sid)
    yield sc, sid
    ctx.release(sid)


@pytest.fixture(scope="module")
def hf_medium(ctx):
    sc = scenes.heightfield_medium(verts=33, sphere_subdiv=2, env_w=256, env_h=128)
    sid = ctx.upload(sc)
    ctx.build_bvh(sid)
    yield sc, sid
    ctx.release(sid)


def render_both(ctx, sid, cam, w, h, **kw):
    out = []
    for flags in (0, abi.OXR_FLAG_NO_CAMERA_CULL):
        ctx.select(sid)
        ctx.resize(w, h)
        ctx.set_camera(cam)
        st = ctx.render(flags=flags, **kw)
        aovs = [ctx.download_aov(i, w, h) for i in range(3)] + [ctx.download_depth(w, h)]
        out.append((aovs, st))
    return out


def assert_same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("filt", [0, 1, 2, 3])
def test_sphere_field_cull_is_exact(ctx, field, filt):
    sc, sid = field
    print(f"[cull] sphere field filter={filt}", flush=True)
    (a, st), (b, st_ref) = render_both(ctx, sid, sc.camera, 160, 90, spp=4, max_depth=6, rr_start=0, filter=filt)
    assert_same(a, b)
    assert st.camera_rays_resolved > 0 and st_ref.camera_rays_resolved == 0
    # every sample's camera ray was either traced or resolved
    assert st.extension_rays + st.camera_rays_resolved == st_ref.extension_rays


@pytest.mark.parametrize("where", ["inside", "away"])
def test_sphere_field_cameras(ctx, field, where):
    sc, sid = field
    print(f"[cull] sphere field camera {where}", flush=True)
    cam = abi.camera((0.5, 0.3, 0.2), (6.0, 0.0, -4.0)) if where == "inside" else abi.camera((0, 0, 30), (0, 0, 60))
    (a, st), (b, _) = render_both(ctx, sid, cam, 96, 64, spp=3, max_depth=6, rr_start=0)
    assert_same(a, b)
    if where == "away":
        assert st.camera_rays_resolved == 96 * 64 * 3 and st.extension_rays == 0


@pytest.mark.parametrize("with_env", [True, False])
def test_lit_sphere_field_cull_is_exact(ctx, with_env):
    # analytic lights in the sky: ruled-out camera rays that meet a sphere or
    # quad light are shaded by the sky pass exactly as the wavefront would
    sc = scenes.lit_sphere_field(with_env=with_env)
    sid = ctx.upload(sc)
    try:
        ctx.build_bvh(sid)
        print(f"[cull] lit sphere field env={with_env}", flush=True)
        (a, st), (b, _) = render_both(ctx, sid, sc.camera, 160, 90, spp=3, max_depth=6, rr_start=0)
        assert_same(a, b)
        assert st.camera_rays_resolved > 0
        assert a[0][..., :3].max() > 0   # the lights are in view
    finally:
        ctx.release(sid)


def test_cornell_nothing_ruled_out(ctx, cornell):
    sc, sid = cornell
    print("[cull] cornell", flush=True)
    (a, st), (b, _) = render_both(ctx, sid, sc.camera, 64, 48, spp=3, max_depth=8, rr_start=3)
    assert_same(a, b)
    assert st.camera_rays_resolved == 0


def test_heightfield_medium_cull_is_exact(ctx, hf_medium):
    sc, sid = hf_medium
    print("[cull] heightfield medium", flush=True)
    (a, _), (b, _) = render_both(ctx, sid, sc.camera, 96, 54, spp=3, max_depth=12, rr_start=0)
    assert_same(a, b)


def test_counting_render_tallies_what_it_would_skip(ctx, field):
    sc, sid = field
    print("[cull] counting render", flush=True)
    ctx.select(sid)
    ctx.resize(160, 90)
    ctx.set_camera(sc.camera)
    st = ctx.render(spp=2, max_depth=6, rr_start=0)
    counted = ctx.render(spp=2, max_depth=6, rr_start=0, flags=abi.OXR_FLAG_COUNT_TRAVERSAL)
    # a counting render traces every camera ray, and tallies the same ones
    assert counted.camera_rays_resolved == st.camera_rays_resolved
    assert counted.extension_rays == st.extension_rays + st.camera_rays_resolved
    assert 0 < counted.resolved_node_fetches < counted.ext_node_fetches
    assert counted.resolved_tri_tests == 0  # a ruled-out ray never reaches a triangle
