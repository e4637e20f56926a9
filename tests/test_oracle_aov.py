"""Pins of the oracle's first-hit AOVs by closed forms.

The depth AOV (an input of the denoiser; the reference has none) is
d / (1 + d) of the distance d from the camera to the first surface along
the camera ray, 1 where the ray leaves the scene, averaged over frames."""
import numpy as np

import oracle_lib
from optixrender_amd import abi, scenes
from optixrender_amd.renderer import SceneData


def _inside_sphere(radius=3.0, subdiv=4):
    tris = scenes.icosphere(subdiv)[:, ::-1] * radius   # reversed winding: faces point inwards
    nrm = -scenes.flat_normals(tris)
    mats = [abi.material(abi.DIFFUSE, albedo=(0.5, 0.5, 0.5), roughness=0.0)]
    cam = abi.camera((0.0, 0.0, 0.0), (0.0, 0.0, -1.0))
    return SceneData(tris.astype(np.float32), nrm.astype(np.float32), np.zeros(len(tris), np.uint32), mats, [], [],
                     None, cam), tris


def test_depth_inside_sphere_is_the_radius():
    sc, tris = _inside_sphere()
    # inscribed radius: distance from the centre to the nearest face plane
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    r_in = np.abs((n * tris[:, 0]).sum(1)).min()
    _, albedo, normal, depth, _ = oracle_lib.render(sc, 24, 16, 3, max_depth=2, with_depth=True)
    lo, hi = r_in / (1 + r_in), 3.0 / 4.0
    assert depth.min() >= lo * (1 - 1e-6) and depth.max() <= hi * (1 + 1e-6)
    assert np.allclose(albedo[..., :3], 0.5) and np.allclose(albedo[..., 3], 1.0)
    # inward-facing flat normals point at the camera: their dot with the view ray is negative
    assert np.all(np.linalg.norm(normal[..., :3], axis=-1) > 0.99)


def test_depth_of_a_miss_is_one_and_means_are_running():
    sc, _ = _inside_sphere()
    sc.camera = abi.camera((0.0, 0.0, 10.0), (0.0, 0.0, 20.0))   # looking away from the sphere
    _, albedo, normal, depth, _ = oracle_lib.render(sc, 8, 8, 2, max_depth=2, with_depth=True)
    assert np.all(depth == 1.0)
    assert np.all(albedo[..., :3] == 0.0) and np.all(normal[..., :3] == 0.0)


def test_depth_progressive_equals_one_call():
    sc = scenes.cornell_spheres()
    c1, a1, n1, d1, _ = oracle_lib.render(sc, 16, 12, 4, max_depth=4, rr_start=3, with_depth=True)
    c2, a2, n2, d2, _ = oracle_lib.render(sc, 16, 12, 2, max_depth=4, rr_start=3, with_depth=True)
    c3, a3, n3, d3, _ = oracle_lib.render(sc, 16, 12, 2, max_depth=4, rr_start=3, frame_base=2, color=c2, albedo=a2,
                                          normal=n2, depth=d2, with_depth=True)
    assert np.array_equal(d1, d3) and np.array_equal(c1, c3)
    assert np.all((d1 > 0) & (d1 < 1))   # the closed box: every camera ray hits a wall or a sphere
