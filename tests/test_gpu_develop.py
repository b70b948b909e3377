"""The output path on the GPU (fl_develop, fl_develop_film) against the
oracle's restatement of color/spectrumoperations.cpp:58-76,215-286: every
colour space x tone map x gamma on random pixel means (negative, zero, large
and NaN included), and a render's resident film developed on the device
against the same film downloaded and developed from the host."""
import itertools

import numpy as np
import pytest

import flgpu
import oracle

pytestmark = pytest.mark.gpu

SPACES = (flgpu.FL_COLOR_SRGB, flgpu.FL_COLOR_ACES2065_1)
TONES = (flgpu.FL_TONEMAP_NONE, flgpu.FL_TONEMAP_ACES_BASIC, flgpu.FL_TONEMAP_REINHARD, flgpu.FL_TONEMAP_FILMIC)
GAMMAS = (flgpu.FL_GAMMA_NONE, flgpu.FL_GAMMA_REC_709, flgpu.FL_GAMMA_1_8, flgpu.FL_GAMMA_2_2, flgpu.FL_GAMMA_4_0)


def inputs():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(0, 1, (4000, 3)), rng.uniform(-0.2, 3, (3000, 3)),
                        np.exp(rng.uniform(-12, 6, (3000, 3))), np.zeros((4, 3)),
                        [[np.nan, 0.5, 0.5], [1e30, 1e30, 1e30]]]).astype(np.float32)
    return x


def test_develop_matches_oracle_every_mode(ctx):
    x = inputs()
    for space, tm, gm in itertools.product(SPACES, TONES, GAMMAS):
        g_rgb, g_u8 = ctx.develop(x, space, tm, gm)
        o_rgb, o_u8 = oracle.develop(x, space, tm, gm)
        # device powf and the host's may differ in the last bits
        np.testing.assert_allclose(g_rgb, o_rgb, rtol=4e-6, atol=1e-7, equal_nan=True,
                                   err_msg="mode %d %d %d" % (space, tm, gm))
        diff = g_u8 != o_u8
        if diff.any():
            # only where the float sits on a rounding boundary of the 8-bit code
            v = np.clip(np.nan_to_num(o_rgb, nan=0.0), 0, 1) * 255 + 0.5
            frac = np.abs(v - np.round(v))
            assert np.all(frac[diff[:, :3]] < 1e-3), (space, tm, gm)


def test_develop_film_equals_host_develop_of_the_film(ctx):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=48, height=40)
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=16, depth=gen.depth, seed=3, flags=flgpu.FL_RENDER_GRAPH)
        s, _, _ = sc.render(p)
        for mode in ((flgpu.FL_COLOR_SRGB, flgpu.FL_TONEMAP_FILMIC, flgpu.FL_GAMMA_2_2),
                     (flgpu.FL_COLOR_ACES2065_1, flgpu.FL_TONEMAP_ACES_BASIC, flgpu.FL_GAMMA_REC_709)):
            f_rgb, f_u8 = ctx.develop_film(0, 48, 40, *mode)
            means = s.reshape(-1, 3) * (np.float32(1) / np.float32(16))
            h_rgb, h_u8 = ctx.develop(means, *mode)
            assert np.array_equal(f_rgb.reshape(-1, 3).view(np.uint32), h_rgb.view(np.uint32))
            assert np.array_equal(f_u8.reshape(-1, 4), h_u8)
            assert f_u8[..., :3].max() > 100                    # a lit image, not zeros


def test_develop_film_of_a_batch_item(ctx):
    gens = [flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=32, height=24),
            flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=40, height=40)]
    scenes = [ctx.scene(g) for g in gens]
    try:
        p = flgpu.default_params(spp=8, depth=8, seed=9)
        films, _ = ctx.render_batch([(sc, None, 8) for sc in scenes], p)
        rgb, _ = ctx.develop_film(1, 40, 40)
        means = films[1][0].reshape(-1, 3) * (np.float32(1) / np.float32(8))
        h_rgb, _ = ctx.develop(means)
        assert np.array_equal(rgb.reshape(-1, 3).view(np.uint32), h_rgb.view(np.uint32))
        with pytest.raises(flgpu.FlError):
            ctx.develop_film(2, 40, 40)
    finally:
        for sc in scenes:
            sc.close()
