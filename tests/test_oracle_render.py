"""The oracle's PathIntegrator restatement pinned by closed forms.

Direct lighting (depth 0: NEE with MIS at the first hit, pathintegrator.cpp:40-43
and integrator.cpp:25-110) of a Lambert floor below a square emitter: the
radiance seen at the point below the emitter's centre is
(R/pi) * E with E = Le * pi * 4 * F_corner (form factor of a parallel
rectangle), so XYZ = to_xyz(R * Le) * 4 F_corner.  The emitter is a primitive
without material, so the camera ray passes through its back face
(pathintegrator.cpp:34-38)."""
import numpy as np
import pytest

import oracle
from scenes import light_over_floor, rect_form_factor_center


@pytest.mark.parametrize("half,height", [(0.5, 1.0), (1.5, 0.8)])
def test_direct_lighting_closed_form(half, height):
    gen = light_over_floor(albedo=(0.5, 0.5, 0.5), half=half, height=height, intensity=5.0)
    sc = oracle.Scene(gen)
    spp = 40000
    s, q, st = sc.render(spp, 0, seed=3, threads=8)
    mean = s[0, 0] / spp
    var = q[0, 0] / spp - mean ** 2
    sigma = np.sqrt(var / spp)
    R = oracle.from_reflectance_rgb((0.5, 0.5, 0.5))
    want = oracle.to_xyz(R * 5.0) * rect_form_factor_center(half, half, height) / np.pi
    assert np.all(np.abs(mean - want) < 4 * sigma + 1e-4 * want), (mean, want, sigma)


def test_pass_through_emitter_seen_from_front_counts_emission():
    # a camera below the emitter looking up sees Le (bounce 0), then passes through
    gen = light_over_floor(half=0.5, height=1.0, intensity=5.0)
    cam = gen.camera
    cam.position[:] = [0.0, 0.5, 0.0]
    cam.look_at[:] = [0.0, 2.0, 0.0]
    cam.up[:] = [0.0, 0.0, 1.0]
    s, _, _ = oracle.Scene(gen).render(64, 0, seed=1, threads=1, camera=cam)
    want = oracle.to_xyz(np.full(60, 5.0, np.float32))
    np.testing.assert_allclose(s[0, 0] / 64, want, rtol=1e-5)
