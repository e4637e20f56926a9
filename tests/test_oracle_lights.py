"""Whole renders of the oracle against closed forms of direct lighting.

A Lambertian floor (Diffuse with roughness 0, Material.h:25-33) lit by one
analytic light and nothing else: a path's bounce 0 adds next-event
estimation (SampleSphere / SampleQuad, Light.h:53-141) and its bounce 1 the
MIS-weighted emission its BSDF ray meets (EvaluateSphere / EvaluateQuad),
after which the ray leaves the scene.  Both strategies together are
unbiased, so the mean radiance of a floor point is albedo/pi times the
irradiance the light gives it:

  sphere of radiance L, radius r, centre at height d, the point at
  horizontal distance x:  E = pi L r^2 d / (d^2 + x^2)^(3/2)
  (the sphere wholly above the horizon);

  rectangle parallel to the floor, the point below one corner, sides a, b
  at height d:  E = L / 2 [A/sqrt(1+A^2) atan(B/sqrt(1+A^2))
                           + B/sqrt(1+B^2) atan(A/sqrt(1+B^2))],
  A = a/d, B = b/d (the corner form factor), summed over four rectangles
  for a point below the centre.

The camera sees a narrow patch of floor around the point and never the
light.  A dropped cosine, a wrong pdf conversion or a missing MIS weight
moves the mean far beyond the 1.5% tolerance."""
import math

import numpy as np
import pytest

import oracle_lib
from optixrender_amd import abi, scenes
from optixrender_amd.renderer import SceneData

ALBEDO = 0.6


def floor_scene(lights):
    q = scenes.quad((-20, 0, -20), (-20, 0, 20), (20, 0, 20), (20, 0, -20))  # normal +y
    pos = q.astype(np.float32)
    nrm = scenes.flat_normals(q).astype(np.float32)
    mats = [abi.material(abi.DIFFUSE, albedo=(ALBEDO, ALBEDO, ALBEDO), roughness=0.0)]
    # a narrow view of the floor around the origin, from the side
    cam = abi.camera((3.0, 1.0, 0.0), (0.0, 0.0, 0.0), cos_fovy=0.02)
    return SceneData(pos, nrm, np.zeros(2, np.uint32), mats, lights, [], None, cam)


def mean_radiance(sc, spp=64):
    col, _, _, _ = oracle_lib.render(sc, 16, 16, spp, max_depth=4, rr_start=8, threads=8)
    return col[..., :3].mean(axis=(0, 1))


def corner_factor(a, b, d):
    A, B = a / d, b / d
    sa, sb = math.sqrt(1 + A * A), math.sqrt(1 + B * B)
    return (A / sa * math.atan(B / sa) + B / sb * math.atan(A / sb)) / (2 * math.pi)


@pytest.mark.parametrize("d,r", [(2.0, 0.5), (1.0, 0.6), (4.0, 0.2)])
def test_sphere_light_direct_lighting(d, r):
    L = 5.0
    sc = floor_scene([abi.light(1, radiance=(L, L, L), position=(0.0, d, 0.0), radius=r)])
    got = mean_radiance(sc)
    want = ALBEDO / math.pi * math.pi * L * r * r / (d * d)   # x ~ 0: the patch is +-0.06 wide
    assert np.allclose(got, want, rtol=0.015), (got, want)


@pytest.mark.parametrize("side,d,double", [(1.0, 1.5, 0), (3.0, 1.0, 1), (0.4, 2.0, 0)])
def test_quad_light_direct_lighting(side, d, double):
    L = 4.0
    h = side / 2
    # u x v = +y: a one-sided quad emits downwards (Light.h:30-34, 59-63)
    lt = abi.light(0, double_side=double, radiance=(L, L, L), position=(-h, d, h), u=(side, 0, 0), v=(0, 0, -side))
    sc = floor_scene([lt])
    got = mean_radiance(sc)
    E = math.pi * L * 4 * corner_factor(h, h, d)
    want = ALBEDO / math.pi * E
    assert np.allclose(got, want, rtol=0.015), (got, want)


def test_one_sided_quad_facing_away_is_dark():
    lt = abi.light(0, double_side=0, radiance=(4, 4, 4), position=(-0.5, 1.5, -0.5), u=(1, 0, 0), v=(0, 0, 1))
    sc = floor_scene([lt])   # u x v = -y: emits upwards only
    assert np.all(mean_radiance(sc, spp=8) == 0.0)


def test_two_lights_keep_the_reference_quirks():
    # With several lights the reference's estimator is biased, and the
    # oracle keeps it so: lightTrace overwrites light_pdf for every light it
    # tests (devicePrograms.cu:192-211), so a BSDF ray that meets a light
    # listed before one it misses gets MIS weight 1 (light_pdf = 0), and the
    # weight of a BSDF hit uses the light's own pdf while next-event
    # estimation divides it by the light count (devicePrograms.cu:429-436).
    # Two visible, non-overlapping lights therefore give more than the sum
    # of each alone, and the total depends on their order.
    sph = abi.light(1, radiance=(5, 5, 5), position=(0.0, 2.0, 0.0), radius=0.5)
    quad = abi.light(0, double_side=0, radiance=(4, 4, 4), position=(-0.5, 1.5, 2.5), u=(1, 0, 0), v=(0, 0, -1))
    a = mean_radiance(floor_scene([sph]))
    b = mean_radiance(floor_scene([quad]))
    sq = mean_radiance(floor_scene([sph, quad]))
    qs = mean_radiance(floor_scene([quad, sph]))
    assert np.all(sq > 1.2 * (a + b)) and np.all(qs > 1.05 * (a + b)), (a, b, sq, qs)
    assert np.all(np.abs(sq - qs) > 0.1 * (a + b)), (sq, qs)
