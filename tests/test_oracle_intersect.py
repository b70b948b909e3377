"""The oracle's hit search, pinned without the oracle's own formulas:

* the watertight test agrees with a float64 Moller-Trumbore solve on hit/miss
  (away from edges) and on t;
* watertightness: rays through the shared edges/vertices of a triangulated
  grid never slip between triangles;
* the SAH BVH returns exactly the brute-force closest hit (bits of t, ID,
  u, v), including the smaller-ID rule on coincident triangles;
* any-hit agrees with closest-hit on occlusion.
"""
import numpy as np

import flgpu
import oracle
from raysets import primary_rays, secondary_rays


def mt64(o, d, a, b, c):
    e1, e2 = b - a, c - a
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-12:
        return None
    s = o - a
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    t = (e2 @ q) / det
    return t, u, v


def test_watertight_matches_float64_solve():
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(4000):
        a, b, c = rng.normal(size=(3, 3))
        o = rng.normal(size=3) * 3
        target = a * rng.random() + b * rng.random() + c * rng.random()
        d = target - o
        d /= np.linalg.norm(d)
        ray = np.array([*o, 0.0, *d, np.inf], np.float32)
        hit, out = oracle.tri_intersect(ray, a, b, c)
        ref = mt64(ray[:3].astype(np.float64), ray[4:7].astype(np.float64), a.astype(np.float32).astype(np.float64),
                   b.astype(np.float32).astype(np.float64), c.astype(np.float32).astype(np.float64))
        if ref is None:
            continue
        t, u, v = ref
        margin = min(u, v, 1 - u - v)
        if abs(margin) < 1e-4 or t <= 1e-3:
            continue
        inside = margin > 0
        assert hit == inside
        if hit:
            assert abs(out[0] - t) <= 1e-5 * abs(t) + 1e-6
            assert abs(out[1] - u) < 1e-4 and abs(out[2] - v) < 1e-4
        checked += 1
    assert checked > 2000


def _grid_scene(n=16):
    xs = np.linspace(-1, 1, n + 1)
    V = np.array([[x, y, 0.0] for y in xs for x in xs], np.float32)
    idx = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            idx += [[a, a + 1, a + n + 2], [a, a + n + 2, a + n + 1]]
    from scenes import lambert
    cam = flgpu.Camera()
    return flgpu.GenSceneData.from_arrays(V, np.array(idx, np.uint32), np.zeros(len(idx), np.uint32),
                                          [lambert((0.5, 0.5, 0.5))], [], np.zeros((1, flgpu.FL_SPD_N)), cam)


def test_watertight_no_leaks_through_shared_edges():
    gen = _grid_scene()
    xs = np.linspace(-1, 1, 17)
    rng = np.random.default_rng(1)
    pts = []
    for x in xs[1:-1]:       # points on shared edges and vertices
        for y in xs[1:-1]:
            pts.append((x, y))
            pts.append((x, y + rng.uniform(-0.06, 0.06)))
            pts.append((x + rng.uniform(-0.06, 0.06), y))
    pts = np.array(pts)
    rays = []
    for (x, y) in pts:
        o = np.array([x, y, 0.0]) + rng.normal(size=3) * np.array([0.3, 0.3, 0]) + np.array([0, 0, 2.0])
        d = np.array([x, y, 0.0]) - o
        d /= np.linalg.norm(d)
        rays.append([*o, 0.0, *d, np.inf])
    rays = np.array(rays, np.float32)
    t, tri, _ = oracle.Scene(gen).intersect(rays, brute=True)
    assert np.all(tri != flgpu.FL_NO_HIT)


def test_bvh_equals_brute_force_bits():
    for cfg, kw in [(flgpu.FL_GEN_CORNELL, {}), (flgpu.FL_GEN_SPHERES, dict(subdiv=2))]:
        gen = flgpu.GenSceneData(cfg, **kw)
        sc = oracle.Scene(gen)
        prim = primary_rays(gen.camera, 96, 96)
        pt, ptri, _ = sc.intersect(prim, brute=True)
        rays = np.concatenate([prim, secondary_rays(gen.triangles(), prim, pt, ptri, 9216)])
        a = sc.intersect(rays)
        b = sc.intersect(rays, brute=True)
        assert np.array_equal(a[1], b[1])
        assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
        assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))


def test_tie_rule_smaller_id_wins():
    from scenes import lambert, quad
    q = quad((0, 0, 0), (1, 0, 0), (0, 1, 0), 1, 1)
    V = np.concatenate([q, q])              # triangles 2,3 duplicate 0,1
    idx = np.arange(12, dtype=np.uint32).reshape(4, 3)
    gen = flgpu.GenSceneData.from_arrays(V, idx[[2, 3, 0, 1]], np.zeros(4, np.uint32), [lambert((1, 1, 1))], [],
                                         np.zeros((1, flgpu.FL_SPD_N)), flgpu.Camera())
    rays = np.array([[0.3, -0.2, 2, 0, 0, 0, -1, np.inf], [-0.4, 0.5, -2, 0, 0, 0, 1, np.inf]], np.float32)
    sc = oracle.Scene(gen)
    for brute in (False, True):
        _, tri, _ = sc.intersect(rays, brute=brute)
        assert set(tri.tolist()) <= {0, 1}


def test_any_hit_matches_closest_hit_occlusion():
    gen = flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=2)
    sc = oracle.Scene(gen)
    prim = primary_rays(gen.camera, 64, 64)
    t, tri, _ = sc.intersect(prim)
    seg = prim.copy()
    seg[:, 7] = np.where(tri != flgpu.FL_NO_HIT, t * 1.001, 3.0)
    _, atri, _ = sc.intersect(seg, flags=flgpu.FL_ANY_HIT)
    assert np.array_equal(atri != flgpu.FL_NO_HIT, tri != flgpu.FL_NO_HIT)
