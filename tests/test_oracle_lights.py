"""Delta illuminants, shadowColor and the NormalIntegrator in the CPU oracle,
pinned by closed forms on a Lambert floor seen from straight above (depth 0:
next-event estimation only):
  point        L = rho/pi * I * cos(theta) / d^2            (point.cpp:12-20)
  spot         the same times the cone falloff (smoothstep) (spot.cpp:23-55)
  directional  L = rho/pi * I * cos(theta)                   (directional.cpp:15-24)
  shadowColor  where the light's sample is occluded, Li = shadowColor
               (integrator.cpp:47-50), so L = rho/pi * shadowColor * cos(theta)
  normals      from_reflectance_rgb(n) of the first hit (normalintegrator.cpp:11-26)
Each pixel's closed form is averaged over an 8 x 8 grid inside the pixel,
as the oracle's jittered camera samples average it."""
import numpy as np

import flgpu
import oracle
from scenes import delta_floor, delta_light, floor_points

ALB = (0.6, 0.5, 0.4)
LAM = 360.0 + np.arange(flgpu.FL_SPD_N, dtype=np.float32)


def spectra():
    warm = (0.6 + 0.4 * np.sin((LAM - 360.0) / 470.0 * np.pi)).astype(np.float32)
    shadow = np.full(flgpu.FL_SPD_N, 0.07, np.float32)
    return np.stack([warm, shadow])


def lit_xyz(spd, rgb=ALB):
    """to_xyz(rho/pi * colour) with the oracle's 60-bin spectra."""
    rho = oracle.from_reflectance_rgb(np.asarray(rgb, np.float32))
    col = oracle.from_spd(LAM, spd)
    return oracle.to_xyz((rho * col / np.float32(np.pi)).astype(np.float32))


def render_oracle(gen, spp=256, integrator=0, with_sigma=False):
    sc = oracle.Scene(gen)
    s, q, _ = sc.render(spp, 0, seed=3, threads=8, integrator=integrator)
    sc.close()
    m = s / spp
    if with_sigma:   # standard error of each pixel mean (from the jittered samples)
        return m, np.sqrt(np.maximum(q / spp - m * m, 0) / (spp - 1))
    return m


def assert_close(img, sigma, want, rel=2e-4):
    """Within 5 standard errors of the pixel mean (the jitter inside a pixel is
    random) plus a small relative allowance for float arithmetic."""
    err = np.abs(img - want)
    bound = 5 * sigma + rel * np.abs(want) + 1e-7
    assert np.all(err <= bound), (float(np.max(err / bound)), int(np.sum(err > bound)))


def spot_falloff(ct, cc, cp):
    t = np.clip((ct - cc) / (cp - cc), 0, 1)
    return np.where(ct < cc, 0.0, np.where(ct > cp, 1.0, t * t * (3 - 2 * t)))


def test_point_light_inverse_square():
    spd = spectra()
    pos = np.array([0.3, 1.5, -0.2])
    gen = delta_floor([delta_light(flgpu.FL_LIGHT_POINT, position=pos, intensity=4.0)], spd)
    img, sig = render_oracle(gen, with_sigma=True)
    P = floor_points(gen.camera, sub=16)
    v = pos - P
    d2 = (v * v).sum(-1)
    g = (4.0 * (v[..., 1] / np.sqrt(d2)) / d2).mean(-1)
    want = g[..., None] * lit_xyz(spd[0])
    assert_close(img, sig, want)


def test_spot_light_cone_and_penumbra():
    spd = spectra()
    pos = np.array([0.0, 1.2, 0.0])
    axis = np.array([0.2, -1.0, 0.1])
    cc, cp = np.cos(np.radians(40.0)), np.cos(np.radians(25.0))
    gen = delta_floor([delta_light(flgpu.FL_LIGHT_SPOT, position=pos, direction=axis, intensity=3.0,
                                   cos_cone=cc, cos_penumbra=cp)], spd)
    img, sig = render_oracle(gen, with_sigma=True)
    P = floor_points(gen.camera, sub=16)
    v = pos - P
    d = np.sqrt((v * v).sum(-1))
    wi = v / d[..., None]
    ct = (-wi * (axis / np.linalg.norm(axis))).sum(-1)
    g = (3.0 * spot_falloff(ct, cc, cp) * wi[..., 1] / d ** 2).mean(-1)
    want = g[..., None] * lit_xyz(spd[0])
    assert_close(img, sig, want)
    # some pixels are dark (outside the cone), some fully lit, some in between
    f = spot_falloff(ct, cc, cp).mean(-1)
    assert (f == 0).sum() > 50 and (f == 1).sum() > 50 and ((f > 0.05) & (f < 0.95)).sum() > 20


def test_directional_light_and_shadow_color():
    spd = spectra()
    direction = np.array([-1.0, -1.0, 0.0])   # travels toward -x and down
    lights = [delta_light(flgpu.FL_LIGHT_DIRECTIONAL, direction=direction, intensity=2.0, shadow_emission=2)]
    gen = delta_floor(lights, spd, occluder=True)
    img, sig = render_oracle(gen, with_sigma=True)
    P = floor_points(gen.camera, sub=16)
    wi = -direction / np.linalg.norm(direction)
    # the slab x in [1.5, 2] at y = 1 shades floor points whose ray toward the
    # light crosses y = 1 inside it: x + 1 in [1.5, 2]
    xs = P[..., 0] + 1.0
    shade = ((xs >= 1.5) & (xs <= 2.0) & (np.abs(P[..., 2]) <= 2.0)).mean(-1)
    lit = lit_xyz(spd[0]) * 2.0 * wi[1]
    dark = lit_xyz(spd[1]) * wi[1]
    want = (1 - shade)[..., None] * lit + shade[..., None] * dark
    assert_close(img, sig, want)
    assert (shade == 1).sum() > 100 and (shade == 0).sum() > 100


def test_area_light_shadow_color_replaces_occluded_light():
    """An area light behind the occluder: its occluded samples add
    rho/pi * shadowColor * cos * w / pdf instead of nothing, so the shadowed
    floor is brighter with a shadow colour than without, and unshadowed
    pixels are the same."""
    spd = spectra()
    area = [((2.6, 2.0, 0.0), 0.3, 5.0, 0)]
    out = []
    for sh in (0, 2):
        gen = delta_floor([], spd, occluder=True, area_lights=area)
        for k in range(gen.n_lights):
            gen.lights[k].shadow_emission = sh
        out.append(render_oracle(gen, spp=256)[..., 1])
    no_sh, with_sh = out
    assert np.all(with_sh >= no_sh - 1e-6)
    assert (with_sh > no_sh + 1e-4).sum() > 20


def test_normal_integrator_is_the_floor_normal_through_the_rgb_basis():
    spd = spectra()
    gen = delta_floor([delta_light(flgpu.FL_LIGHT_POINT, position=(0, 1, 0))], spd)
    img = render_oracle(gen, spp=4, integrator=1)
    want = oracle.to_xyz(oracle.from_reflectance_rgb(np.array([0.0, 1.0, 0.0], np.float32)))
    np.testing.assert_allclose(img, np.broadcast_to(want, img.shape), rtol=1e-6, atol=1e-7)
    # a normal with negative components goes through the clamp (-W + M <= 0 here)
    s = oracle.from_reflectance_rgb(np.array([0.0, -1.0, 0.0], np.float32))
    assert np.all(s >= 0)
