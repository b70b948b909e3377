"""Image textures on the GPU (ImageMapTexture, shading/textures/imagemap.cpp:38-62):
the textured floors of test_oracle_textures.py against the same closed form
(Lambert kd, StandardSurface kd, with and without per-vertex uvs), and a
Cornell box of standard surfaces whose kd, ks, kt, sheen and coat colours are
image textures against the oracle's render at equal spp (per-pixel z-test as
in test_gpu_render.py, parity mode)."""
import numpy as np
import pytest

import flgpu
import oracle
from scenes import checker_texture, textured_floor
from test_gpu_lights import assert_within, render_gpu
from test_gpu_render import z_stats
from test_gpu_standard import standard_box
from test_oracle_lights import spectra
from test_oracle_textures import want_image

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("standard,with_uvs", [(False, True), (True, True), (False, False)])
def test_textured_floor_closed_form(ctx, standard, with_uvs):
    tex = checker_texture()
    gen = textured_floor(tex, spectra(), with_uvs=with_uvs, standard=standard)
    img, sig, _, _ = render_gpu(ctx, gen, 256)
    want, _ = want_image(gen, tex, 2.0, with_uvs)
    assert_within(img, sig, want)


def textured_box(width=32, height=32):
    """standard_box with image textures for kd (walls), ks, sheen, coat and kt
    (blocks), and per-vertex uvs from a planar map of the positions."""
    gen = standard_box(width, height)
    gen.textures = [checker_texture(7, 5, seed=s) for s in range(5)]
    ss = gen.standard_surfaces
    ss[0].kd_texture, ss[0].sheen_texture, ss[0].coat_texture = 1, 2, 3
    ss[1].kd_texture = 4
    ss[2].ks_texture = 5
    ss[3].kd_texture, ss[3].kt_texture = 2, 3
    v = gen.vertices
    gen.uvs = np.stack([(v[:, 0] + v[:, 2] + 2.0) / 4.0, (v[:, 1] + 1.0) / 2.0], -1).astype(np.float32)
    return gen


def test_textured_box_consistent_with_oracle(ctx):
    gen = textured_box()
    spp = 1024
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=spp, depth=gen.depth, mode=flgpu.FL_SPECTRAL_PARITY, seed=61)
        gs, gq, st = sc.render(p)
    assert st.nan_samples == 0
    os_, oq, _ = oracle.Scene(gen).render(spp, gen.depth, seed=62, threads=16)
    frac, allowed, mean_z, bad = z_stats(gs, gq, os_, oq, spp)
    print("textured box: frac %.4f allowed %.4f mean z %.2f" % (frac, allowed, mean_z))
    assert bad == 0 and frac <= allowed and abs(mean_z) < 3
    # the textures change the image
    plain = standard_box()
    ps, _, _ = oracle.Scene(plain).render(64, plain.depth, seed=63, threads=16)
    assert abs(os_.mean() / spp - ps.mean() / 64) > 0.02 * (ps.mean() / 64)


def test_bad_texture_index_is_rejected(ctx):
    gen = textured_floor(checker_texture(), spectra())
    gen.materials[0].texture = 2          # only one texture
    with pytest.raises(flgpu.FlError):
        ctx.scene(gen)


def test_batch_with_textured_and_plain_scenes(ctx):
    """A batch mixing a textured scene, a standard-surface scene and a plain
    Lambert scene: each item's film equals rendering it alone with its seed
    (the plain one alone runs the smaller shade kernel)."""
    gens = [textured_box(24, 20), flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=20, height=16), standard_box(16, 16)]
    seed = 77
    flags = flgpu.FL_RENDER_GRAPH
    scenes = [ctx.scene(g) for g in gens]
    try:
        p = flgpu.default_params(spp=8, depth=8, seed=seed, flags=flags)
        films, st = ctx.render_batch([(sc, None, 8) for sc in scenes], p)
        assert st.nan_samples == 0
        for k, sc in enumerate(scenes):
            ps = flgpu.default_params(spp=8, depth=8, seed=(seed + k * flgpu.ITEM_SEED_STRIDE) & ((1 << 64) - 1),
                                      flags=flags)
            single, _, _ = sc.render(ps)
            assert np.array_equal(films[k][0].view(np.uint32), single.view(np.uint32)), k
    finally:
        for sc in scenes:
            sc.close()
