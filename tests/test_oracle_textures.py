"""Material textures in the oracle: the RGBA8 lookup (tex2D<float4> with
linear filtering, wrap addressing and normalised-float reads,
Renderer.cpp:95-144), NormalFromTangentToWorld (Utils.h:85-94) and the
overrides of __closesthit__radiance (devicePrograms.cu:86-115).

Pins: closed forms of bilinear interpolation (texel centres, constant and
piecewise-linear images, wrap periodicity, channel order, row order), an
independent float64 bilinear over random images, hand-derived tangent
frames, rotation invariants of the normal map, and whole renders in which
a texture that reads exactly 0 must equal the untextured material."""
import math

import numpy as np
import pytest

import oracle_lib as orc
from optixrender_amd import abi, scenes


def bilinear64(img, u, v):
    """Textbook bilinear interpolation between the four nearest texel centres
    (texel (x, y) centred at ((x + 0.5) / w, (y + 0.5) / h)), periodic."""
    h, w = img.shape[:2]
    x = (u % 1.0) * w - 0.5
    y = (v % 1.0) * h - 0.5
    x0, y0 = math.floor(x), math.floor(y)
    ax, ay = x - x0, y - y0
    f = img.astype(np.float64) / 255.0
    g = lambda i, j: f[j % h, i % w]
    return (g(x0, y0) * (1 - ax) * (1 - ay) + g(x0 + 1, y0) * ax * (1 - ay) + g(x0, y0 + 1) * (1 - ax) * ay
            + g(x0 + 1, y0 + 1) * ax * ay)


def test_channel_order_and_normalisation():
    img = np.zeros((3, 5, 4), np.uint8)
    img[...] = (10, 20, 30, 40)
    got = orc.texture_sample(img, 0.37, 0.61)
    # byte / 255 per channel, R first (a constant image: every weight sums to 1)
    assert np.allclose(got, np.array([10, 20, 30, 40]) / 255.0, rtol=0, atol=2e-7)
    # 0 and 255 read as exactly 0 and (to an ulp) 1
    img[...] = (0, 255, 0, 255)
    got = orc.texture_sample(img, 0.1, 0.9)
    assert got[0] == 0.0 and got[2] == 0.0 and abs(got[1] - 1.0) < 2e-7


def test_texel_centres_and_row_order():
    rng = np.random.default_rng(1)
    h, w = 8, 16  # powers of two: texel-centre coordinates are exact floats
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    for y in range(h):
        for x in range(w):
            got = orc.texture_sample(img, (x + 0.5) / w, (y + 0.5) / h)
            # row y sits at v = (y + 0.5) / h: row 0 is at v ~ 0
            assert np.array_equal(got, img[y, x].astype(np.float32) / np.float32(255.0))


def test_linear_between_centres():
    # an image that is linear in x between texel centres is reproduced exactly
    # by bilinear filtering (away from the wrap seam)
    w, h = 16, 4
    img = np.zeros((h, w, 4), np.uint8)
    img[..., 0] = (np.arange(w) * 16)[None, :]
    for u in np.linspace(0.5 / w, (w - 1.5) / w, 37):
        got = orc.texture_sample(img, float(u), 0.3)
        want = (u * w - 0.5) * 16 / 255.0
        assert abs(got[0] - want) < 1e-5, u


def test_wrap_is_periodic():
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (5, 7, 4), dtype=np.uint8)
    for u, v in [(0.125, 0.375), (0.5, 0.0625), (0.96875, 0.03125)]:  # exact binary fractions
        ref = orc.texture_sample(img, u, v)
        for du, dv in [(1, 0), (0, 1), (-1, 0), (3, -2), (-7, 5)]:
            assert np.array_equal(orc.texture_sample(img, u + du, v + dv), ref)
    # the seam interpolates between the last and the first texel
    img2 = np.zeros((1, 4, 4), np.uint8)
    img2[0, 0, 0] = 200
    got = orc.texture_sample(img2, 0.0, 0.5)     # halfway between texel 3 and texel 0
    assert abs(got[0] - 100 / 255.0) < 1e-6


def test_non_finite_coordinates_read_at_zero():
    rng = np.random.default_rng(6)
    img = rng.integers(0, 256, (5, 7, 4), dtype=np.uint8)
    ref = orc.texture_sample(img, 0.0, 0.25)
    for bad in (float("nan"), float("inf"), -float("inf")):
        assert np.array_equal(orc.texture_sample(img, bad, 0.25), ref)
    assert np.array_equal(orc.texture_sample(img, 0.5, float("inf")), orc.texture_sample(img, 0.5, 0.0))


def test_matches_float64_bilinear_on_random_images():
    rng = np.random.default_rng(3)
    for h, w in [(1, 1), (1, 9), (13, 1), (29, 37), (64, 48)]:
        img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
        for _ in range(60):
            u, v = rng.uniform(-3, 3, 2)
            got = orc.texture_sample(img, float(np.float32(u)), float(np.float32(v)))
            want = bilinear64(img, float(np.float32(u)), float(np.float32(v)))
            assert np.allclose(got, want, atol=3e-6), (h, w, u, v)


def test_normal_from_tangent_frames():
    # n = +z: up = +x (|n.z| >= 0.9), TangentX = cross(x, z) = -y,
    # TangentY = cross(z, -y) = +x
    z = [0.0, 0.0, 1.0]
    assert np.allclose(orc.normal_from_tangent(z, [1.0, 0.5, 0.5]), [0, -1, 0], atol=1e-7)
    assert np.allclose(orc.normal_from_tangent(z, [0.5, 1.0, 0.5]), [1, 0, 0], atol=1e-7)
    assert np.allclose(orc.normal_from_tangent(z, [0.5, 0.5, 1.0]), [0, 0, 1], atol=1e-7)
    # n = +x: up = +z, TangentX = cross(z, x) = +y, TangentY = cross(x, y) = +z
    x = [1.0, 0.0, 0.0]
    assert np.allclose(orc.normal_from_tangent(x, [1.0, 0.5, 0.5]), [0, 1, 0], atol=1e-7)
    assert np.allclose(orc.normal_from_tangent(x, [0.5, 1.0, 0.5]), [0, 0, 1], atol=1e-7)
    assert np.allclose(orc.normal_from_tangent(x, [0.5, 0.5, 1.0]), [1, 0, 0], atol=1e-7)
    # a texel value below 0.5 points the other way
    assert np.allclose(orc.normal_from_tangent(x, [0.0, 0.5, 0.5]), [0, -1, 0], atol=1e-7)


def test_normal_from_tangent_is_a_rotation():
    rng = np.random.default_rng(4)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        t1, t2 = rng.uniform(0, 1, (2, 3))
        a = orc.normal_from_tangent(n, t1).astype(np.float64)
        b = orc.normal_from_tangent(n, t2).astype(np.float64)
        e1 = (2 * t1 - 1) / np.linalg.norm(2 * t1 - 1)
        e2 = (2 * t2 - 1) / np.linalg.norm(2 * t2 - 1)
        assert abs(np.linalg.norm(a) - 1) < 1e-6
        assert abs(np.dot(a, n) - e1[2]) < 1e-5       # z of the tangent normal is along n
        assert abs(np.dot(a, b) - np.dot(e1, e2)) < 1e-5  # angles are kept
        # right-handed frame: det [tx ty n] = +1
        tx = orc.normal_from_tangent(n, [1.0, 0.5, 0.5]).astype(np.float64)
        ty = orc.normal_from_tangent(n, [0.5, 1.0, 0.5]).astype(np.float64)
        assert abs(np.linalg.det(np.stack([tx, ty, n])) - 1) < 1e-5


@pytest.mark.parametrize("slot,field,value", [
    (0, "albedo", (0.0, 0.0, 0.0)),
    (4, "metallic", 0.0),
])
def test_zero_texture_equals_zero_parameter(slot, field, value):
    # a texture that reads exactly 0 everywhere replaces its slot's parameter
    # by 0; the render must equal the one with that parameter set to 0 (the
    # right sphere of the Cornell scene made metal-workflow for the test)
    def scene(**kw):
        sc = scenes.cornell_spheres(subdiv=1)
        args = dict(roughness=0.3, albedo=(0.8, 0.6, 0.4), metallic=0.7)
        args.update(kw)
        sc.materials[4] = abi.material(abi.METAL_WORKFLOW, **args)
        return sc

    def render(sc):
        return orc.render(sc, 24, 24, 2, max_depth=6, rr_start=3, threads=4)

    plain = scene(**{field: value})
    tex = scene()
    tex.textures = [np.zeros((3, 5, 4), np.uint8)]
    tex.materials[4].tex[slot] = 0
    for x, y in zip(render(plain)[:3], render(tex)[:3]):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    # images that no material names change nothing
    unused = scene()
    unused.textures = [np.full((2, 2, 4), 255, np.uint8)]
    assert np.array_equal(render(scene())[0].view(np.uint32), render(unused)[0].view(np.uint32))


def test_textured_scene_renders_and_textures_matter():
    sc = scenes.textured_cornell(subdiv=1)
    col, alb, nrm, _ = orc.render(sc, 32, 32, 2, max_depth=6, rr_start=3, threads=4)
    assert np.isfinite(col).all()
    flat = scenes.textured_cornell(subdiv=1)
    for m in flat.materials:
        m.tex[:] = [-1] * 8
    col2, alb2, nrm2, _ = orc.render(flat, 32, 32, 2, max_depth=6, rr_start=3, threads=4)
    # the albedo checker shows in the first-hit albedo; normal maps in the normal AOV
    assert np.abs(alb - alb2).max() > 0.1
    assert np.abs(nrm - nrm2).max() > 0.05
    # without coordinates every triangle samples (0,0), (1,0), (0,1): still textured
    nouv = scenes.textured_cornell(subdiv=1, with_uvs=False)
    col3, alb3, _, _ = orc.render(nouv, 32, 32, 2, max_depth=6, rr_start=3, threads=4)
    assert np.isfinite(col3).all() and np.abs(alb3 - alb).max() > 0.01
