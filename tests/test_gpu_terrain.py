"""Spectral dispersion and the terrain workload (fBm heightfield, Cauchy
prisms, GGX metals, 8 cameras, guided) at reduced size.

Dispersion is not in the reference (scalar eta). With a wavelength-dependent
eta only the hero wavelength can follow a refracted path: the secondary
lanes end at the dispersive event and the hero lane's weight is multiplied
by 4, which keeps the estimate unbiased. Checked by rendering glass with a
negligible Cauchy B (the dispersive code path, physically the same glass)
against B = 0: the images must agree statistically."""
import numpy as np
import pytest

import flgpu
from test_gpu_render import z_stats

pytestmark = pytest.mark.gpu
FLAGS = flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH


def glass_cornell(cauchy_b):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=40, height=40)
    m = gen.materials[3]                  # the two boxes
    m.kind = flgpu.FL_MAT_DIELECTRIC
    m.ior = 1.5
    m.cauchy_b = cauchy_b
    m.rgb[:] = [1.0, 1.0, 1.0]
    m.rgb_t[:] = [1.0, 1.0, 1.0]
    m.weight = 1.0
    return gen


def test_dispersion_termination_is_unbiased(ctx):
    spp = 1024
    with ctx.scene(glass_cornell(0.0)) as a, ctx.scene(glass_cornell(1e-7)) as b:
        sa, qa, st_a = a.render(flgpu.default_params(spp=spp, depth=8, seed=31, flags=FLAGS))
        sb, qb, st_b = b.render(flgpu.default_params(spp=spp, depth=8, seed=32, flags=FLAGS))
    assert st_a.nan_samples == 0 and st_b.nan_samples == 0
    frac, allowed, mean_z, bad = z_stats(sb, qb, sa, qa, spp)
    print("dispersive vs scalar glass: frac |z|>3 %.4f (allowed %.4f), mean z %.2f" % (frac, allowed, mean_z))
    assert bad == 0 and frac <= allowed and abs(mean_z) < 3


def test_dispersion_changes_the_image(ctx):
    """With the config's B = 0.0042 um^2 the refracted light splits by
    wavelength: the X/Z balance of the image moves while luminance stays."""
    spp = 256
    with ctx.scene(glass_cornell(0.0)) as a, ctx.scene(glass_cornell(0.0042)) as b:
        sa, _, _ = a.render(flgpu.default_params(spp=spp, depth=8, seed=5, flags=FLAGS))
        sb, _, _ = b.render(flgpu.default_params(spp=spp, depth=8, seed=5, flags=FLAGS))
    d = np.abs(sa - sb).sum() / np.abs(sa).sum()
    assert d > 1e-3
    assert abs(sa[..., 1].mean() - sb[..., 1].mean()) / sa[..., 1].mean() < 0.05


def test_terrain_reduced_guided_cameras_consistent(ctx):
    """The terrain config at reduced size: 256 x 128 heightfield, 200 Cauchy
    prisms, 24 GGX metals, 8 cameras rendered as one batch. Guided (trained
    online) and plain batches agree statistically, camera by camera pooled."""
    W, H, spp = 64, 36, 256
    gen = flgpu.GenSceneData(flgpu.FL_GEN_TERRAIN, grid_x=256, grid_y=128, width=W, height=H)
    cams = [flgpu.GenSceneData(flgpu.FL_GEN_TERRAIN, grid_x=2, grid_y=2, camera_index=k, width=W, height=H).camera
            for k in range(8)]
    with ctx.scene(gen) as sc, ctx.guide(seed=7) as g:
        p = flgpu.default_params(spp=spp, depth=gen.depth, seed=41, flags=FLAGS)
        plain, st_p = ctx.render_batch([(sc, c, spp) for c in cams], p)
        pg = flgpu.default_params(spp=spp, depth=gen.depth, seed=42,
                                  flags=FLAGS | flgpu.FL_RENDER_GUIDED | flgpu.FL_RENDER_TRAIN_GUIDE)
        pg.guide = g.h
        guided, st_g = ctx.render_batch([(sc, c, spp) for c in cams], pg)
        steps, _, div = g.stats()
    assert st_p.nan_samples == 0 and st_g.nan_samples == 0 and div == 0 and steps > 0
    gs = np.concatenate([f[0] for f in guided])
    gq = np.concatenate([f[1] for f in guided])
    ps = np.concatenate([f[0] for f in plain])
    pq = np.concatenate([f[1] for f in plain])
    frac, allowed, mean_z, bad = z_stats(gs, gq, ps, pq, spp)
    print("terrain guided vs plain (8 cameras): frac |z|>3 %.4f (allowed %.4f), mean z %.2f" % (frac, allowed, mean_z))
    assert bad == 0 and frac <= allowed and abs(mean_z) < 3
