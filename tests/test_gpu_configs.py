"""The batched, guided and terrain workloads at their full stated sizes.

- 64-scene batch (triangle counts log-uniform in [1e3, 2e6], 128^2 to
  1024^2, 16 spp, depth 12), rendered as one fl_render_batch in parity mode:
  every item's film equals its single render bit for bit, and the smallest
  item agrees with the oracle (PathIntegrator::Li,
  integrators/pathintegrator.cpp:12-87) in its image mean and in 16 x 16-pixel
  block means.
- Guided room (the spheres scene in a closed room lit through a 5% slot,
  2.6 M triangles, 1024^2, depth 16): guided (trained online) and plain
  renders agree in block means and image mean.
- Terrain (4096 x 2048 fBm heightfield, 16.8 M triangles, 200 Cauchy prisms,
  24 GGX metals, 8 cameras at 3840 x 2160, depth 24): a guided batch and a
  plain batch at a few spp agree in block means, with no NaN sample.

Block means pool many pixels, so their z-scores are close to normal even
at a few samples per pixel; a block's variance is the sum of its pixels'
sample variances of the mean. Pass: at most 0.27% + 3 binomial sigmas of the
blocks beyond |z| = 3, and the image-mean |z| < 4 per channel."""
import numpy as np
import pytest

import flgpu
import oracle

pytestmark = pytest.mark.gpu
MASK = (1 << 64) - 1
NO_RR = 1 << 20


def block_stats(s, q, n, bh, bw):
    """Per-block mean of pixel means and its variance, (H/bh, W/bw, 3) each
    (the image cropped to whole blocks)."""
    H, W = s.shape[0] // bh * bh, s.shape[1] // bw * bw
    s = s[:H, :W].astype(np.float64)
    q = q[:H, :W].astype(np.float64)
    m = s / n
    var = np.maximum(q / n - m * m, 0.0) * n / (n - 1) / n
    P = bh * bw
    mb = m.reshape(H // bh, bh, W // bw, bw, 3).sum((1, 3)) / P
    vb = var.reshape(H // bh, bh, W // bw, bw, 3).sum((1, 3)) / (P * P)
    return mb, vb


def compare_blocks(a, b, what):
    """a, b: lists of (block means, block variances) of two renders of the
    same views. Asserts the block and image-mean consistency rule."""
    ma = np.concatenate([x[0].reshape(-1, 3) for x in a])
    va = np.concatenate([x[1].reshape(-1, 3) for x in a])
    mb = np.concatenate([x[0].reshape(-1, 3) for x in b])
    vb = np.concatenate([x[1].reshape(-1, 3) for x in b])
    den = np.sqrt(va + vb)
    valid = den[:, 1] > 0
    z = (ma[valid, 1] - mb[valid, 1]) / den[valid, 1]
    nb = int(valid.sum())
    frac = float(np.mean(np.abs(z) > 3))
    allowed = 0.0027 + 3 * np.sqrt(0.0027 * 0.9973 / max(nb, 1))
    zi = (ma.mean(0) - mb.mean(0)) / np.sqrt(va.sum(0) + vb.sum(0)) * ma.shape[0]
    print("%s: %d blocks, frac |z|>3 %.4f (allowed %.4f), image-mean z %s" % (what, nb, frac, allowed, np.round(zi, 2)))
    assert frac <= allowed, (what, frac, allowed)
    assert np.all(np.abs(zi) < 4), (what, zi)


def test_batch_64_scenes_full_size(ctx):
    gens = [flgpu.GenSceneData(flgpu.FL_GEN_BATCH, scene_index=k) for k in range(64)]
    spp, depth, seed = 16, 12, 77
    assert all(g.spp == spp for g in gens)
    tris = [g.n_triangles for g in gens]
    assert min(tris) >= 1000 and max(tris) <= 2_000_000
    assert {g.camera.width for g in gens} <= {128, 256, 512, 1024}
    scenes = [ctx.scene(g) for g in gens]
    try:
        p = flgpu.default_params(spp=spp, depth=depth, mode=flgpu.FL_SPECTRAL_PARITY, seed=seed,
                                 flags=flgpu.FL_RENDER_GRAPH)
        films, st = ctx.render_batch([(sc, None, spp) for sc in scenes], p)
        assert st.nan_samples == 0
        assert st.samples == sum(g.camera.width * g.camera.height * spp for g in gens)
        for k, sc in enumerate(scenes):
            ps = flgpu.default_params(spp=spp, depth=depth, mode=flgpu.FL_SPECTRAL_PARITY,
                                      seed=(seed + k * flgpu.ITEM_SEED_STRIDE) & MASK, flags=flgpu.FL_RENDER_GRAPH)
            single, sq, _ = sc.render(ps)
            assert np.array_equal(films[k][0].view(np.uint32), single.view(np.uint32)), k
            assert np.array_equal(films[k][1].view(np.uint32), sq.view(np.uint32)), k
        # the smallest item against the oracle (its own seed stream)
        k = int(np.argmin([t * g.camera.width * g.camera.height for t, g in zip(tris, gens)]))
        g = gens[k]
        os_, oq, _ = oracle.Scene(g).render(spp, depth, seed=5, threads=16)
        gs, gq = films[k]
        compare_blocks([block_stats(gs, gq, spp, 16, 16)], [block_stats(os_, oq, spp, 16, 16)],
                       "batch item %d (%d tris, %dx%d) vs oracle" % (k, tris[k], g.camera.width, g.camera.height))
    finally:
        for sc in scenes:
            sc.close()


def test_guided_room_full_size_unbiased(ctx):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_GUIDED)
    assert gen.camera.width == 1024 and gen.n_triangles > 2_600_000
    spp = 128
    flags = flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH
    with ctx.scene(gen) as sc, ctx.guide(seed=1) as g:
        p = flgpu.default_params(spp=spp, depth=gen.depth, seed=3, flags=flags)
        p.rr_start = NO_RR
        ps, pq, st_p = sc.render(p)
        pg = flgpu.default_params(spp=spp, depth=gen.depth, seed=4,
                                  flags=flags | flgpu.FL_RENDER_GUIDED | flgpu.FL_RENDER_TRAIN_GUIDE)
        pg.rr_start = NO_RR
        pg.guide = g.h
        gs, gq, st_g = sc.render(pg)
        steps, _, div = g.stats()
    assert st_p.nan_samples == 0 and st_g.nan_samples == 0 and div == 0 and steps > 0
    compare_blocks([block_stats(gs, gq, spp, 32, 32)], [block_stats(ps, pq, spp, 32, 32)],
                   "guided room 1024^2 guided vs plain")


def test_terrain_full_size_guided_batch(ctx):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_TERRAIN)
    assert gen.n_triangles >= 16_764_930
    cams = [flgpu.GenSceneData(flgpu.FL_GEN_TERRAIN, grid_x=2, grid_y=2, camera_index=k, width=3840,
                               height=2160).camera for k in range(8)]
    spp = 4
    flags = flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH
    with ctx.scene(gen) as sc, ctx.guide(seed=7) as g:
        p = flgpu.default_params(spp=spp, depth=gen.depth, seed=41, flags=flags)
        plain, st_p = ctx.render_batch([(sc, c, spp) for c in cams], p)
        bp = [block_stats(s, q, spp, 60, 60) for s, q in plain]
        del plain
        pg = flgpu.default_params(spp=spp, depth=gen.depth, seed=42,
                                  flags=flags | flgpu.FL_RENDER_GUIDED | flgpu.FL_RENDER_TRAIN_GUIDE)
        pg.guide = g.h
        guided, st_g = ctx.render_batch([(sc, c, spp) for c in cams], pg)
        bg = [block_stats(s, q, spp, 60, 60) for s, q in guided]
        del guided
        steps, _, div = g.stats()
    assert st_p.nan_samples == 0 and st_g.nan_samples == 0 and div == 0 and steps > 0
    assert st_p.samples == 8 * 3840 * 2160 * spp
    compare_blocks(bg, bp, "terrain 8 x 3840x2160 guided vs plain")
