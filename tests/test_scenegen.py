"""Procedural workloads: triangle counts, orientation, determinism."""
import numpy as np

import flgpu


def normals(gen):
    t = gen.triangles().astype(np.float64)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    return n / np.linalg.norm(n, axis=1, keepdims=True), t.mean(1)


def test_cornell_counts_and_orientation():
    g = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL)
    assert g.n_triangles == 36 and g.n_lights == 2
    n, c = normals(g)
    # walls face the box interior
    assert np.all(np.sum(n[:10] * -c[:10], axis=1) > 0)
    # light faces down and is emissive
    lt = [g.lights[i].tri for i in range(g.n_lights)]
    assert np.allclose(n[lt], [0, -1, 0])
    assert (g.camera.width, g.camera.height, g.spp, g.depth) == (256, 256, 16, 8)


def test_spheres_counts_outward_normals_and_determinism():
    g = flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=3)
    assert g.n_triangles == 8 * 20 * 4 ** 3 + 4
    n, c = normals(g)
    sph = slice(2, 2 + 8 * 1280)
    mats = g.tri_material[sph]
    # per sphere, normals point away from the sphere's centroid
    for k in range(8):
        s = slice(2 + k * 1280, 2 + (k + 1) * 1280)
        ctr = c[s].mean(0)
        assert np.mean(np.sum(n[s] * (c[s] - ctr), axis=1) > 0) > 0.99
    g2 = flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=3)
    assert np.array_equal(g.vertices, g2.vertices) and np.array_equal(mats, g2.tri_material[sph])


def test_full_config1_triangle_count():
    g = flgpu.GenSceneData(flgpu.FL_GEN_SPHERES)
    assert g.n_triangles == 8 * 327680 + 2 + 2
    assert (g.camera.width, g.spp, g.depth) == (1024, 64, 16)


def test_batch_scenes_cover_the_triangle_range():
    counts = [flgpu.GenSceneData(flgpu.FL_GEN_BATCH, scene_index=i).n_triangles for i in range(0, 64, 8)]
    assert min(counts) >= 500 and max(counts) <= 3_000_000


def test_guided_room_slot_is_five_percent():
    g = flgpu.GenSceneData(flgpu.FL_GEN_GUIDED, subdiv=1)
    t = g.triangles().astype(np.float64)
    n, c = normals(g)
    area = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
    ceiling = (np.abs(c[:, 1] - 6.0) < 1e-4) & np.all(np.isclose(n, [0, -1, 0], atol=1e-6), axis=1)
    assert abs(area[ceiling].sum() - 0.95 * 96.0) < 1e-3
