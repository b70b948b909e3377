"""The autotuner (fl_autotune): probe renders choose BVH node width,
traversal, queue and shading knobs. It must return a valid configuration no
slower than the default, and the knobs it sets other than the pass size must
not change the image (they only reorder work): every node width gives the
same hits and the same film, bit for bit."""
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
        assert r.node_width in (4, 6, 8)
        fresh = flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=5, width=192, height=128)
        fresh.bvh_width = r.node_width
        with ctx.scene(fresh) as sc2:
            assert sc2.info().n_nodes == sc.info().n_nodes   # the scene was left at the chosen width
        assert r.stack_depth in (6, 8, 12, 16) and r.trav_budget in (8, 16, 32)
        base, _, _ = sc.render(p)
        q = flgpu.default_params(spp=16, depth=gen.depth, seed=4, flags=p.flags)
        r.apply(q)
        q.max_paths = 0   # same pass structure as the default render
        tuned, _, st = sc.render(q)
        assert st.nan_samples == 0
    assert np.array_equal(base.view(np.uint32), tuned.view(np.uint32))


def test_node_widths_give_identical_hits_and_films(ctx):
    from raysets import primary_rays, secondary_rays

    films, hits, nodes = [], [], []
    rays = None
    for w in (8, 6, 4, 2):
        gen = flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=3, width=64, height=48)
        gen.bvh_width = w
        with ctx.scene(gen) as sc:
            nodes.append(sc.info().n_nodes)
            if rays is None:   # primary rays and cosine bounces off their hits
                prim = primary_rays(gen.camera, 128, 96)
                t0, tri0, _ = sc.intersect(prim)
                rays = np.concatenate([prim, secondary_rays(gen.triangles(), prim, t0, tri0, 20000)])
            t, tri, uv = sc.intersect(rays)
            hits.append((t.view(np.uint32).copy(), tri.copy(), uv.view(np.uint32).copy()))
            f, _, st = sc.render(flgpu.default_params(spp=8, depth=gen.depth, seed=2, flags=flgpu.FL_RENDER_GRAPH))
            assert st.nan_samples == 0
            films.append(f)
    # binary nodes make the most nodes (between the wider ones the greedy
    # collapse of a restructured tree need not be monotonic)
    assert nodes[3] > max(nodes[:3]), nodes
    assert (hits[0][1] != flgpu.FL_NO_HIT).mean() > 0.3
    for k in range(1, 4):
        for a, b in zip(hits[0], hits[k]):
            assert np.array_equal(a, b), k
        assert np.array_equal(films[0].view(np.uint32), films[k].view(np.uint32)), k


def test_bad_node_width_is_rejected(ctx):
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=8, height=8)
    gen.bvh_width = 9
    with pytest.raises(flgpu.FlError):
        ctx.scene(gen)


def test_traversal_refill_modes_give_identical_films(ctx):
    """trav_prefetch 1 (next rays loaded ahead), 2 (budgeted refill) and 3
    (one ray per thread), at several refill thresholds, only reorder the
    traversal: on a small scene (auto: one ray per thread) and a larger one
    (auto: budgeted refill for this small pass) every film equals the
    automatic choice's bit for bit. The persistent modes end each launch by
    splitting the last rays' traversals across a wave's free lanes
    (fl_trace.h trace_drain); one ray per thread never does, so the
    comparison also checks that the split parts merge to the same hits."""
    for kind, kw in ((flgpu.FL_GEN_CORNELL, dict(width=96, height=96)),
                     (flgpu.FL_GEN_SPHERES, dict(subdiv=5, width=128, height=96))):
        gen = flgpu.GenSceneData(kind, **kw)
        with ctx.scene(gen) as sc:
            films = []
            for mode, lanes in ((0, 0), (1, 16), (1, 1), (1, 64), (2, 16), (2, 1), (3, 0)):
                p = flgpu.default_params(spp=8, depth=gen.depth, seed=5, flags=flgpu.FL_RENDER_GRAPH)
                p.trav_prefetch = mode
                p.refill_lanes = lanes
                f, _, st = sc.render(p)
                assert st.nan_samples == 0
                films.append(f)
            for k in range(1, len(films)):
                assert np.array_equal(films[0].view(np.uint32), films[k].view(np.uint32)), (kind, k)
            p = flgpu.default_params(spp=1, depth=gen.depth, seed=5)
            p.trav_prefetch = 4
            with pytest.raises(flgpu.FlError):
                sc.render(p)
