"""Delta illuminants, shadowColor and the NormalIntegrator on the GPU:
closed-form direct lighting from point, spot and directional lights on a
Lambert floor (parity mode, depth 0), a box lit by an area light and three
delta lights with shadow colours against the CPU oracle at equal spp
(per-pixel z-test as in test_gpu_render.py), the normal image against the
oracle, and argument checks of fl_scene_create."""
import ctypes as C

import numpy as np
import pytest

import flgpu
import oracle
from scenes import delta_floor, delta_light, floor_points
from test_gpu_render import z_stats
from test_oracle_lights import lit_xyz, spectra, spot_falloff

pytestmark = pytest.mark.gpu

PARITY = flgpu.FL_SPECTRAL_PARITY


def render_gpu(ctx, gen, spp, depth=0, flags=0, seed=5):
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=spp, depth=depth, mode=PARITY, seed=seed, flags=flags)
        s, q, st = sc.render(p)
    assert st.nan_samples == 0
    m = s / spp
    return m, np.sqrt(np.maximum(q / spp - m * m, 0) / (spp - 1)), s, q


def assert_within(img, sigma, want):
    err = np.abs(img - want)
    bound = 5 * sigma + 2e-4 * np.abs(want) + 1e-7
    assert np.all(err <= bound), (float(np.max(err / bound)), int(np.sum(err > bound)))


def test_point_spot_directional_closed_form(ctx):
    spd = spectra()
    P = None
    # point
    pos = np.array([0.3, 1.5, -0.2])
    gen = delta_floor([delta_light(flgpu.FL_LIGHT_POINT, position=pos, intensity=4.0)], spd)
    img, sig, _, _ = render_gpu(ctx, gen, 512)
    P = floor_points(gen.camera, sub=16)
    v = pos - P
    d2 = (v * v).sum(-1)
    want = (4.0 * (v[..., 1] / np.sqrt(d2)) / d2).mean(-1)[..., None] * lit_xyz(spd[0])
    assert_within(img, sig, want)
    # spot
    axis = np.array([0.2, -1.0, 0.1])
    spos = np.array([0.0, 1.2, 0.0])
    cc, cp = np.cos(np.radians(40.0)), np.cos(np.radians(25.0))
    gen = delta_floor([delta_light(flgpu.FL_LIGHT_SPOT, position=spos, direction=axis, intensity=3.0,
                                   cos_cone=cc, cos_penumbra=cp)], spd)
    img, sig, _, _ = render_gpu(ctx, gen, 512)
    v = spos - P
    d = np.sqrt((v * v).sum(-1))
    wi = v / d[..., None]
    ct = (-wi * (axis / np.linalg.norm(axis))).sum(-1)
    want = (3.0 * spot_falloff(ct, cc, cp) * wi[..., 1] / d ** 2).mean(-1)[..., None] * lit_xyz(spd[0])
    assert_within(img, sig, want)
    # directional with an occluder and a shadow colour
    direction = np.array([-1.0, -1.0, 0.0])
    gen = delta_floor([delta_light(flgpu.FL_LIGHT_DIRECTIONAL, direction=direction, intensity=2.0,
                                   shadow_emission=2)], spd, occluder=True)
    img, sig, _, _ = render_gpu(ctx, gen, 512)
    wd = -direction / np.linalg.norm(direction)
    xs = P[..., 0] + 1.0
    shade = ((xs >= 1.5) & (xs <= 2.0) & (np.abs(P[..., 2]) <= 2.0)).mean(-1)
    want = (1 - shade)[..., None] * lit_xyz(spd[0]) * 2.0 * wd[1] + shade[..., None] * lit_xyz(spd[1]) * wd[1]
    assert_within(img, sig, want)


def lit_box(width=32, height=32):
    """The Cornell box with its area light (now with a shadow colour), a
    point light with a shadow colour, a spot aimed at the floor and a
    directional light shining in through the open front."""
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=width, height=height)
    shadow = np.full(flgpu.FL_SPD_N, 0.05, np.float32)
    gen.emission = np.concatenate([gen.emission, shadow])
    gen.n_emission += 1
    sh = gen.n_emission          # 1 + index of the shadow spectrum
    for k in range(gen.n_lights):
        gen.lights[k].shadow_emission = sh
    dl = [delta_light(flgpu.FL_LIGHT_POINT, position=(-0.5, 0.6, 0.4), intensity=0.8, shadow_emission=sh),
          delta_light(flgpu.FL_LIGHT_SPOT, position=(0.4, 0.9, 0.5), direction=(0.0, -1.0, -0.3), intensity=1.5,
                      cos_cone=np.cos(np.radians(35)), cos_penumbra=np.cos(np.radians(20))),
          delta_light(flgpu.FL_LIGHT_DIRECTIONAL, direction=(0.3, -0.4, -1.0), intensity=0.6)]
    gen.delta_lights = (flgpu.DeltaLight * len(dl))(*dl)
    gen.n_delta_lights = len(dl)
    return gen


def test_lit_box_consistent_with_oracle(ctx):
    gen = lit_box()
    spp = 1024
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=spp, depth=gen.depth, mode=PARITY, seed=31)
        gs, gq, st = sc.render(p)
    assert st.nan_samples == 0
    os_, oq, _ = oracle.Scene(gen).render(spp, gen.depth, seed=32, threads=16)
    frac, allowed, mean_z, bad = z_stats(gs, gq, os_, oq, spp)
    print("lit box: frac %.4f allowed %.4f mean z %.2f" % (frac, allowed, mean_z))
    assert bad == 0 and frac <= allowed and abs(mean_z) < 3
    # the delta lights and shadow colours change the image: it is not the plain box
    plain = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=32, height=32)
    ps, _, _ = oracle.Scene(plain).render(64, plain.depth, seed=33, threads=16)
    assert abs(os_.mean() / spp - ps.mean() / 64) > 0.05 * (ps.mean() / 64)


def test_normal_integrator_consistent_with_oracle(ctx):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=40, height=40)
    spp = 256
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=spp, depth=gen.depth, mode=PARITY, seed=41, flags=flgpu.FL_RENDER_NORMALS)
        gs, gq, st = sc.render(p)
    assert st.nan_samples == 0
    assert st.rays_shadow == 0                      # no light transport at all
    os_, oq, _ = oracle.Scene(gen).render(spp, gen.depth, seed=42, threads=16, integrator=1)
    frac, allowed, mean_z, bad = z_stats(gs, gq, os_, oq, spp)
    print("normals: frac %.4f allowed %.4f mean z %.2f" % (frac, allowed, mean_z))
    assert bad == 0 and frac <= allowed and abs(mean_z) < 3


def test_bad_delta_lights_are_rejected(ctx):
    spd = spectra()
    bad = [delta_light(7, position=(0, 1, 0)),                                   # unknown type
           delta_light(flgpu.FL_LIGHT_POINT, position=(0, 1, 0), emission=5),    # no such emission
           delta_light(flgpu.FL_LIGHT_SPOT, position=(0, 1, 0)),                 # spot without axis
           delta_light(flgpu.FL_LIGHT_POINT, position=(0, 1, 0), shadow_emission=9)]
    for d in bad:
        gen = delta_floor([d], spd)
        with pytest.raises(flgpu.FlError):
            ctx.scene(gen)
