"""The GPU path loop (k_shade's depth breaks, roulette, etaScale, emission
after specular bounces, NEE + MIS with the continuation as the BSDF sample)
against the closed forms of tests/furnace.py and against the oracle on the
same furnaces, in parity mode (and hero mode for one depth).
PathIntegrator::Li: integrators/pathintegrator.cpp:12-87."""
import numpy as np
import pytest

import flgpu
import oracle
from furnace import (LE, expected_glass_parity, expected_hero, expected_parity, furnace, glass_pixels, image_mean_z,
                     parity_xyz, spd_bins, two_sample_z)
from test_gpu_render import z_stats
from test_tables import load_tables

pytestmark = pytest.mark.gpu

NO_RR = 1 << 20
Z_OK = 4.5
GLASS_DEPTHS = (64, 64, 64)


def gpu_render(ctx, gen, spp, depths, rr_start, mode=flgpu.FL_SPECTRAL_PARITY, rr_threshold=1.0, seed=5):
    p = flgpu.default_params(spp=spp, depth=depths[0], mode=mode, seed=seed)
    p.diffuse_depth, p.reflection_depth, p.transmission_depth = depths
    p.rr_start = rr_start
    p.rr_threshold = rr_threshold
    with ctx.scene(gen) as sc:
        s, q, st = sc.render(p)
    assert st.nan_samples == 0
    return s, q


@pytest.mark.parametrize("D,rr_start", [(0, NO_RR), (1, NO_RR), (3, NO_RR), (8, NO_RR), (2, 0), (3, 0), (8, 0), (8, 3)])
def test_lambert_furnace_closed_form_and_oracle(ctx, D, rr_start):
    rho, spp = 0.8, 4096
    gen = furnace(rho)
    s, q = gpu_render(ctx, gen, spp, (D, D, D), rr_start)
    exp = expected_parity(rho, D)
    z, mean, sigma = image_mean_z(s, q, spp, exp)
    print("D=%d rr_start=%d: z %s, relative sigma %.2e" % (D, rr_start, np.round(z, 2), sigma[1] / mean[1]))
    assert np.all(np.abs(z) < Z_OK), (z, mean, exp)
    for Dn in (D - 1, D + 1):
        if Dn >= 0:
            assert np.all(np.abs((mean - expected_parity(rho, Dn)) / sigma) > 8.0), Dn
    ospp = 256 if D >= 8 else 128
    os_, oq, _ = oracle.Scene(gen).render(ospp, D, seed=17, threads=16, rr_start=rr_start)
    _, om, osg = image_mean_z(os_, oq, ospp, exp)
    zz = two_sample_z((mean, sigma), (om, osg))
    assert np.all(np.abs(zz) < Z_OK), zz


def test_lambert_furnace_hero_closed_form(ctx):
    rho, spp, D = 0.8, 4096, 3
    s, q = gpu_render(ctx, furnace(rho), spp, (D, D, D), 0, mode=flgpu.FL_SPECTRAL_HERO)
    exp = expected_hero(rho, D, load_tables())
    z, mean, _ = image_mean_z(s, q, spp, exp)
    assert np.all(np.abs(z) < Z_OK), (z, mean, exp)


def test_glass_furnace_closed_form_and_oracle(ctx):
    rho, spp = 0.5, 4096
    gen = furnace(rho, glass=True)
    inside, outside = glass_pixels(gen)
    s, q = gpu_render(ctx, gen, spp, GLASS_DEPTHS, 3)
    exp = expected_glass_parity(rho)
    os_, oq, _ = oracle.Scene(gen).render(128, 0, seed=17, threads=16, rr_start=3, depths=GLASS_DEPTHS)
    for mask in (inside, outside, None):
        z, mean, sigma = image_mean_z(s, q, spp, exp, mask)
        assert np.all(np.abs(z) < Z_OK), (z, mean, exp)
        _, om, osg = image_mean_z(os_, oq, 128, exp, mask)
        assert np.all(np.abs(two_sample_z((mean, sigma), (om, osg))) < Z_OK)


def test_eta_scale_keeps_lossless_glass_at_the_spectral_noise(ctx):
    """Black emitting walls: every sample through lossless glass equals, lane
    for lane, what a direct view of the wall gives (beta * etaScale stays 1
    inside the glass, so roulette below threshold 0.99 never fires). The
    glass pixels' relative sample variance is then the hero-wavelength noise
    alone, as in the pixels that see the wall directly."""
    spp = 4096
    gen = furnace(0.0, glass=True)
    inside, outside = glass_pixels(gen)
    s, q = gpu_render(ctx, gen, spp, GLASS_DEPTHS, 0, rr_threshold=0.99)
    want = parity_xyz(spd_bins(1.0) * LE)
    m = s.astype(np.float64) / spp
    rel_var = np.maximum(q.astype(np.float64) / spp - m * m, 0.0) / np.maximum(m * m, 1e-30)
    for mask in (inside, outside):
        z, mean, _ = image_mean_z(s, q, spp, want, mask)
        assert np.all(np.abs(z) < Z_OK), (z, mean, want)
    ratio = rel_var[inside].mean(0) / rel_var[outside].mean(0)
    print("relative variance glass / direct:", ratio)
    assert np.all((ratio > 0.9) & (ratio < 1.1)), ratio


def test_standard_stack_with_specular_lobe_furnace(ctx):
    """Every wall a standard surface mixing diffuse transmission, Fresnel
    specular and clearcoat, and every wall an emitter: light sampling and the
    continuation's MIS weights must partition unity for the stack's
    non-specular directions (the continuation draws them with density
    S/(m+k) of m non-specular and k specular lobes; estimate_direct's BSDF
    half with S/m, integrator.cpp:63-66)."""
    from stdsurf import material
    gen = furnace(0.8, W=24, H=24)
    gen.standard_surfaces = (flgpu.StandardSurface * 1)(material("difftrans+fspec+coat"))
    gen.n_standard_surfaces = 1
    gen.materials[0].kind = flgpu.FL_MAT_STANDARD
    gen.materials[0].standard = 0
    spp, D = 1024, 3
    s, q = gpu_render(ctx, gen, spp, (D, D, D), 3)
    os_, oq, _ = oracle.Scene(gen).render(spp, D, seed=12, threads=16, rr_start=3)
    frac, allowed, mean_z, bad = z_stats(s, q, os_, oq, spp)
    print("frac |z|>3 = %.4f (allowed %.4f), image-mean z = %.2f" % (frac, allowed, mean_z))
    assert bad == 0 and frac <= allowed and abs(mean_z) < 3
