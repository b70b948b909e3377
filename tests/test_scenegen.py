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


def test_skylight_room_is_twelve_triangles_with_a_five_percent_strip():
    g = flgpu.GenSceneData(flgpu.FL_GEN_GUIDED_SKYLIGHT, subdiv=1)
    t = g.triangles().astype(np.float64)
    n, c = normals(g)
    area = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
    lt = [g.lights[i].tri for i in range(g.n_lights)]
    # the room: the 12 triangles after the spheres and ground, before the light
    room = np.arange(lt[0] - 12, lt[0])
    wall_mat = g.tri_material[room[0]]
    assert np.all(g.tri_material[room] == wall_mat)
    # every room face looks into the room [-6,6] x [0,6] x [-4,4]
    centre = np.array([0.0, 3.0, 0.0])
    assert np.all(np.sum(n[room] * (centre - c[room]), axis=1) > 0)
    ceiling = room[np.abs(c[room, 1] - 6.0) < 1e-4]
    assert len(ceiling) == 2 and abs(area[ceiling].sum() - 0.95 * 96.0) < 1e-3
    # the strip left open is z in [-4, -3.6]; the light above it faces down
    assert np.isclose(t[ceiling][:, :, 2].min(), -3.6, atol=1e-5)
    assert np.allclose(n[lt], [0, -1, 0]) and np.all(t[lt][:, :, 1] > 6.0)
    lz = t[lt][:, :, 2]
    assert lz.min() < -4.0 and lz.max() > -3.6
