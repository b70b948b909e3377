"""Batched rendering (fl_render_batch): items of different scenes, sizes and
sample counts share one wavefront (one persistent launch per kernel stage,
one graph-replayed bounce loop). Every item's film must equal the film of
rendering that item alone with the item's seed, bit for bit: per-path
results do not depend on which other paths share the queues."""
import numpy as np
import pytest

import flgpu

pytestmark = pytest.mark.gpu
MASK = (1 << 64) - 1


def test_batch_items_equal_single_renders(ctx):
    gens = [flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=40, height=28),
            flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=2, width=33, height=21),
            flgpu.GenSceneData(flgpu.FL_GEN_BATCH, scene_index=3, subdiv=-1, width=24, height=24),
            flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=17, height=9)]
    spps = [16, 8, 12, 16]
    seed = 123
    flags = flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH
    scenes = [ctx.scene(g) for g in gens]
    try:
        p = flgpu.default_params(spp=16, depth=8, seed=seed, flags=flags)
        films, st = ctx.render_batch([(sc, None, spp) for sc, spp in zip(scenes, spps)], p)
        assert st.samples == sum(g.camera.width * g.camera.height * s for g, s in zip(gens, spps))
        assert st.nan_samples == 0
        for k, (sc, spp) in enumerate(zip(scenes, spps)):
            ps = flgpu.default_params(spp=spp, depth=8, seed=(seed + k * flgpu.ITEM_SEED_STRIDE) & MASK, flags=flags)
            single, _, _ = sc.render(ps)
            assert np.array_equal(films[k][0].view(np.uint32), single.view(np.uint32)), k
    finally:
        for sc in scenes:
            sc.close()


def test_batch_of_one_equals_render(ctx):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=32, height=32)
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=8, depth=8, seed=9, flags=flgpu.FL_RENDER_GRAPH)
        films, _ = ctx.render_batch([(sc, None, 0)], p)
        single, _, _ = sc.render(p)
    assert np.array_equal(films[0][0].view(np.uint32), single.view(np.uint32))
