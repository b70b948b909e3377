"""Structural invariants of the GPU-built 8-wide quantized BVH (the role of
Embree's build behind BVH::initialize_scene_geometry/finalise,
bvh/bvh.cpp:7-47), read back through fl_scene_download_bvh:

- every triangle record is reached from the root exactly once, through
  exactly one leaf slot, and the records' IDs are a permutation of the input
  triangles, each record carrying that triangle's vertices, material and
  light flag;
- every dequantized slot box (org + q 2^e, exact in fp32) contains the
  triangles of its leaf, and every used slot of an internal child's node lies
  inside the parent slot's dequantized box;
- every node is reached exactly once, leaf slots hold 1..3 triangles, and a
  node has at most `bvh_width` used slots;
- every node's grid origin is a multiple of its step (so dequantization is
  exact).
Checked at widths 8, 6, 4 and 2 on the Cornell box and on the spheres scene at
subdivision 5 (163,842 triangles). The checker is itself shown to reject a
slot box shrunk off its triangles and a triangle listed twice."""
import numpy as np
import pytest

import flgpu

pytestmark = pytest.mark.gpu

NODE = np.dtype([("org", "<f4", 3), ("e", "u1", 3), ("imask", "u1"), ("child_base", "<u4"), ("tri_base", "<u4"),
                 ("meta", "u1", 8), ("qlo", "<u2", (3, 8)), ("qhi", "<u2", (3, 8))])
TRI = np.dtype([("v0", "<f4", 3), ("id", "<u4"), ("v1", "<f4", 3), ("material", "<u4"), ("v2", "<f4", 3),
                ("light", "<u4")])
assert NODE.itemsize == 128 and TRI.itemsize == 48


class BvhError(AssertionError):
    pass


def check(nodes_raw, tris_raw, gen, width):
    nodes = nodes_raw.view(NODE).ravel()
    tris = tris_raw.view(TRI).ravel()
    n_nodes, n_tris = len(nodes), len(tris)
    step = np.ldexp(1.0, nodes["e"].astype(np.int64) - 127)                  # (n, 3), exact in float64
    org = nodes["org"].astype(np.float64)
    if np.any(np.mod(org, step) != 0):
        raise BvhError("a node origin is not a multiple of its step")
    lo = org[:, :, None] + nodes["qlo"].astype(np.float64) * step[:, :, None]   # (n, 3, 8)
    hi = org[:, :, None] + nodes["qhi"].astype(np.float64) * step[:, :, None]
    imask = nodes["imask"].astype(np.int64)
    meta = nodes["meta"].astype(np.int64)
    internal = ((imask[:, None] >> np.arange(8)) & 1).astype(bool)             # (n, 8)
    leaf = ~internal & (meta != 0)
    used = internal | leaf
    if np.any(used.sum(1) > width):
        raise BvhError("a node has more used slots than the node width")
    cnt = meta >> 5
    if np.any((cnt[leaf] < 1) | (cnt[leaf] > 3)):
        raise BvhError("a leaf slot does not hold 1..3 triangles")
    # internal children: child_base + rank among the node's internal slots
    rank = np.cumsum(internal, 1) - internal
    child = nodes["child_base"][:, None].astype(np.int64) + rank
    node_seen = np.zeros(n_nodes, np.int64)
    tri_seen = np.zeros(n_tris, np.int64)
    node_seen[0] = 1
    frontier = np.array([0])
    while frontier.size:
        f_int = internal[frontier]
        parents, slots = np.nonzero(f_int)
        pn = frontier[parents]
        kids = child[pn, slots]
        if kids.size and (kids.max() >= n_nodes):
            raise BvhError("a child index beyond the node array")
        np.add.at(node_seen, kids, 1)
        # every used slot of each child inside the parent slot's box
        plo, phi = lo[pn, :, slots], hi[pn, :, slots]                           # (m, 3)
        ku = used[kids]                                                         # (m, 8)
        klo, khi = lo[kids], hi[kids]                                           # (m, 3, 8)
        inside = (klo >= plo[:, :, None]) & (khi <= phi[:, :, None])
        if np.any(ku & ~inside.all(1)):
            raise BvhError("a child box is not inside its parent slot's box")
        # leaves of the frontier: triangles inside their slot box
        ln, ls = np.nonzero(leaf[frontier])
        fn = frontier[ln]
        first = nodes["tri_base"][fn].astype(np.int64) + (meta[fn, ls] & 31)
        for j in range(3):
            sel = cnt[fn, ls] > j
            r = first[sel] + j
            if r.size and r.max() >= n_tris:
                raise BvhError("a triangle record index beyond the record array")
            np.add.at(tri_seen, r, 1)
            blo, bhi = lo[fn[sel], :, ls[sel]], hi[fn[sel], :, ls[sel]]
            for v in ("v0", "v1", "v2"):
                p = tris[v][r].astype(np.float64)
                if np.any((p < blo) | (p > bhi)):
                    raise BvhError("a triangle is outside its leaf slot's box")
        frontier = kids
    if np.any(node_seen != 1):
        raise BvhError("a node is reached %s times" % sorted(set(node_seen[node_seen != 1].tolist()))[:3])
    if np.any(tri_seen != 1):
        raise BvhError("a triangle record is in %s leaves" % sorted(set(tri_seen[tri_seen != 1].tolist()))[:3])
    ids = tris["id"].astype(np.int64)
    if not np.array_equal(np.sort(ids), np.arange(gen.n_triangles)):
        raise BvhError("record IDs are not a permutation of the triangles")
    V = gen.triangles()
    for j, v in enumerate(("v0", "v1", "v2")):
        if not np.array_equal(tris[v], V[ids, j]):
            raise BvhError("a record's vertices differ from its triangle's")
    mats = np.asarray(gen.tri_material, np.uint32)[ids]
    kinds = np.array([m.kind for m in gen.materials])
    none = (mats == 0xffff) | (kinds[np.minimum(mats, len(kinds) - 1)] == flgpu.FL_MAT_NONE)
    if not np.array_equal(np.where(none, 0xffff, mats), tris["material"]):
        raise BvhError("a record's material differs from its triangle's")
    light = np.zeros(gen.n_triangles, np.uint32)
    for l in range(gen.n_lights):
        light[gen.lights[l].tri] = l + 1
    if not np.array_equal(light[ids], tris["light"]):
        raise BvhError("a record's light flag differs from its triangle's")
    return nodes, tris


SCENES = {
    "cornell": lambda: flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=16, height=16),
    "spheres5": lambda: flgpu.GenSceneData(flgpu.FL_GEN_SPHERES, subdiv=5, width=16, height=16),
}


@pytest.fixture(scope="module")
def scenes():
    return {k: f() for k, f in SCENES.items()}


@pytest.mark.parametrize("name", list(SCENES))
@pytest.mark.parametrize("width", [8, 6, 4, 2])
def test_bvh_invariants(ctx, scenes, name, width):
    gen = scenes[name]
    gen.bvh_width = width
    try:
        with ctx.scene(gen) as sc:
            info = sc.info()
            nodes_raw, tris_raw = sc.bvh()
    finally:
        gen.bvh_width = 0
    assert info.n_triangles == gen.n_triangles and info.n_leaf_tris == gen.n_triangles
    nodes, _ = check(nodes_raw, tris_raw, gen, width)
    assert len(nodes) == info.n_nodes


def test_checker_rejects_broken_trees(ctx, scenes):
    gen = scenes["spheres5"]
    with ctx.scene(gen) as sc:
        nodes_raw, tris_raw = sc.bvh()
    check(nodes_raw, tris_raw, gen, 8)
    # the first leaf slot of the tree (in node order), shrunk to nothing
    nodes = nodes_raw.copy().view(NODE).ravel()
    n = None
    for i in range(len(nodes)):
        leafs = [s for s in range(8) if not (nodes["imask"][i] >> s) & 1 and nodes["meta"][i][s]]
        if leafs:
            n, s = i, leafs[0]
            break
    nodes["qhi"][n][:, s] = nodes["qlo"][n][:, s]
    with pytest.raises(BvhError, match="outside its leaf"):
        check(nodes.view(np.uint8).reshape(-1, 128), tris_raw, gen, 8)
    # a leaf slot pointing at its neighbour's triangles: one record twice
    nodes = nodes_raw.copy().view(NODE).ravel()
    m = int(nodes["meta"][n][s])
    nodes["meta"][n][s] = (m & 0xE0) | (((m & 31) + 1) & 31)
    with pytest.raises(BvhError):
        check(nodes.view(np.uint8).reshape(-1, 128), tris_raw, gen, 8)
