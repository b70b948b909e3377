"""Image statistics of the GPU render path against the CPU restatement of
PathIntegrator::Li (integrators/pathintegrator.cpp:12-87) at equal spp.

Parity mode samples wavelengths uniformly in 400..700 nm and looks every
spectrum up in the reference's 60 bins, so its expected XYZ equals the
reference's to_xyz(Li) (color/spectrumoperations.cpp:193-207).  Consistency:
per-pixel z = (m_gpu - m_cpu) / sqrt(s2_gpu/n + s2_cpu/n); the fraction of
|z| > 3 must stay within 0.27% + 3 binomial sigmas, and the image-mean z must
satisfy |z| < 3."""
import numpy as np
import pytest

import flgpu
import oracle

pytestmark = pytest.mark.gpu


def z_stats(gs, gq, os_, oq, n):
    gm, om = gs / n, os_ / n
    gv = np.maximum(gq / n - gm * gm, 0) * n / (n - 1)
    ov = np.maximum(oq / n - om * om, 0) * n / (n - 1)
    den = np.sqrt(gv / n + ov / n)
    valid = den > 0
    z = np.zeros_like(gm)
    z[valid] = (gm[valid] - om[valid]) / den[valid]
    both_zero = (~valid) & (gm == om)
    frac = float(np.mean(np.abs(z[valid]) > 3)) if valid.any() else 0.0
    npx = int(valid.sum())
    allowed = 0.0027 + 3 * np.sqrt(0.0027 * 0.9973 / max(npx, 1))
    # image mean of Y: pixels are independent, so variances add (X, Y, Z of one
    # pixel are strongly correlated, so they are not pooled)
    y = (Ellipsis, 1)
    mean_z = float((gm[y].mean() - om[y].mean()) / np.sqrt(np.sum(gv[y] / n + ov[y] / n)) * gm[y].size)
    return frac, allowed, mean_z, int((~valid & ~both_zero).sum())


@pytest.mark.parametrize("config", [flgpu.FL_GEN_CORNELL])
def test_parity_image_consistent_with_oracle(ctx, config):
    # 46x44: edge tiles of the 8x8 path layout are clipped
    gen = flgpu.GenSceneData(config, width=46, height=44)
    spp = 1024
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=spp, depth=gen.depth, mode=flgpu.FL_SPECTRAL_PARITY, seed=11)
        gs, gq, st = sc.render(p)
    assert st.nan_samples == 0
    os_, oq, ost = oracle.Scene(gen).render(spp, gen.depth, seed=12, threads=16)
    frac, allowed, mean_z, bad = z_stats(gs, gq, os_, oq, spp)
    print("frac |z|>3 = %.4f (allowed %.4f), image-mean z = %.2f, cpu %.1fs" % (frac, allowed, mean_z, ost.seconds))
    assert bad == 0
    assert frac <= allowed
    assert abs(mean_z) < 3


def test_spheres_parity_consistent(ctx):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=3, width=32, height=32)
    spp = 1024
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=spp, depth=gen.depth, mode=flgpu.FL_SPECTRAL_PARITY, seed=21)
        gs, gq, st = sc.render(p)
    os_, oq, _ = oracle.Scene(gen).render(spp, gen.depth, seed=22, threads=16)
    frac, allowed, mean_z, bad = z_stats(gs, gq, os_, oq, spp)
    print("spheres: frac %.4f allowed %.4f mean z %.2f" % (frac, allowed, mean_z))
    assert bad == 0 and frac <= allowed and abs(mean_z) < 3


@pytest.mark.parametrize("config", [flgpu.FL_GEN_CORNELL, flgpu.FL_GEN_SPHERES])
def test_one_light_nee_consistent_with_oracle(ctx, config):
    """FL_RENDER_NEE_ONE: one uniformly chosen light per vertex instead of
    uniform_sample_all_lights (integrators/integrator.cpp:10-23). The
    expectation is the same, so the film passes the z-test against the
    oracle, which samples every light; the two-triangle emitter of both
    scenes gets half the shadow rays."""
    kw = {"subdiv": 3, "width": 32, "height": 32} if config == flgpu.FL_GEN_SPHERES else {"width": 46, "height": 44}
    gen = flgpu.GenSceneData(config, **kw)
    spp = 1024
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=spp, depth=gen.depth, mode=flgpu.FL_SPECTRAL_PARITY, seed=31,
                                 flags=flgpu.FL_RENDER_NEE_ONE)
        gs, gq, st = sc.render(p)
        _, _, st_all = sc.render(flgpu.default_params(spp=spp, depth=gen.depth, mode=flgpu.FL_SPECTRAL_PARITY,
                                                      seed=31), want_film=False)
    assert st.nan_samples == 0
    print("shadow rays: one light %d, all lights %d" % (st.rays_shadow, st_all.rays_shadow))
    assert st.rays_shadow < 0.6 * st_all.rays_shadow
    os_, oq, _ = oracle.Scene(gen).render(spp, gen.depth, seed=32, threads=16)
    frac, allowed, mean_z, bad = z_stats(gs, gq, os_, oq, spp)
    print("one light: frac %.4f allowed %.4f mean z %.2f" % (frac, allowed, mean_z))
    assert bad == 0 and frac <= allowed and abs(mean_z) < 3


def test_render_deterministic(ctx):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=64, height=64)
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=8, depth=8, seed=5)
        a, _, _ = sc.render(p)
        b, _, _ = sc.render(p)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_bounce_loop_variants_identical(ctx):
    """The pipelined loop, the captured-graph loop (replayed many times over
    several passes and two calls), the eager per-bounce loop with host
    reads of the queue sizes, and shading with and without the per-block
    material sort give bit-identical films."""
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=60, height=44)
    npix = 64 * 48   # 8x8 tiles, padded
    films = {}
    with ctx.scene(gen) as sc:
        for name, flags in [("eager", flgpu.FL_RENDER_TIMING), ("pipelined", 0), ("graph", flgpu.FL_RENDER_GRAPH),
                            ("graph again", flgpu.FL_RENDER_GRAPH),
                            ("graph counters", flgpu.FL_RENDER_GRAPH | flgpu.FL_RENDER_COUNTERS),
                            ("sorted shading", flgpu.FL_RENDER_GRAPH | flgpu.FL_RENDER_SORT_SHADING),
                            ("eager again", flgpu.FL_RENDER_TIMING)]:
            p = flgpu.default_params(spp=8, depth=8, seed=9, flags=flags)
            p.max_paths = 2 * npix   # 2 samples per pixel per pass: 4 passes
            f, _, st = sc.render(p)
            assert st.passes == 4 and st.nan_samples == 0
            assert st.rays_extension >= 60 * 44 * 8   # every camera ray was queued
            films[name] = f
    ref = films["eager"].view(np.uint32)
    for name, f in films.items():
        assert np.array_equal(f.view(np.uint32), ref), name


def test_hero_mode_close_to_parity_mode(ctx):
    """Hero (360..830 nm, continuous spectra) and parity (60 bins) estimate
    slightly different integrals; their image means agree to a few percent."""
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=64, height=64)
    with ctx.scene(gen) as sc:
        a, _, _ = sc.render(flgpu.default_params(spp=256, depth=8, mode=flgpu.FL_SPECTRAL_HERO, seed=1))
        b, _, _ = sc.render(flgpu.default_params(spp=256, depth=8, mode=flgpu.FL_SPECTRAL_PARITY, seed=2))
    ma, mb = a.reshape(-1, 3).mean(0), b.reshape(-1, 3).mean(0)
    assert np.all(np.abs(ma - mb) / mb < 0.05), (ma, mb)
