"""The autotuner (fl_autotune): probe renders choose traversal, queue and
shading knobs. It must return a valid configuration no slower than the
default, and the knobs it sets other than the pass size must not change the
image (they only reorder work)."""
import numpy as np
import pytest

import flgpu

pytestmark = pytest.mark.gpu


def test_autotune_valid_and_image_preserving(ctx):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=5, width=192, height=128)
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=16, depth=gen.depth, seed=4, flags=flgpu.FL_RENDER_VNDF_PDF | flgpu.FL_RENDER_GRAPH)
        r = ctx.autotune(sc, p)
        print("tuned: max_paths %d stack %d waves %d budget %d shade_waves %d sort %d | %.2f ms vs default %.2f ms, %d probes" % (
            r.max_paths, r.stack_depth, r.trav_waves, r.trav_budget, r.shade_waves, r.sort_shading, r.ms_best,
            r.ms_default, r.n_tried))
        assert r.n_tried >= 8
        assert 0 < r.ms_best <= r.ms_default
        assert r.stack_depth in (6, 8, 12, 16) and r.trav_budget in (8, 16, 32)
        base, _, _ = sc.render(p)
        q = flgpu.default_params(spp=16, depth=gen.depth, seed=4, flags=p.flags)
        r.apply(q)
        q.max_paths = 0   # same pass structure as the default render
        tuned, _, st = sc.render(q)
        assert st.nan_samples == 0
    assert np.array_equal(base.view(np.uint32), tuned.view(np.uint32))
