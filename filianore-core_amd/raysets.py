"""Ray sets of the 4M-ray intersection workload: coherent primary rays and cosine-diffuse
secondary rays spawned from primary hits (spawn rule of
core/interaction.h:11-31 / :60-63: offset along the geometric normal by
|n|.pError, pError = |p|*gamma(5), tMin = 1e-4)."""
import numpy as np

GAMMA5 = np.float32((5 * 5.9604644775390625e-08) / (1 - 5 * 5.9604644775390625e-08))


def camera_basis(cam):
    pos = np.array(cam.position, np.float64)
    la = np.array(cam.look_at, np.float64)
    up = np.array(cam.up, np.float64)
    fd = np.linalg.norm(pos - la)
    hh = np.tan(np.radians(cam.vfov_deg) / 2)
    hw = cam.width / cam.height * hh
    fwd = (la - pos) / fd
    u = np.cross(up, fwd)
    u /= np.linalg.norm(u)
    v = np.cross(fwd, u)
    llc = pos - u * hw * fd - v * hh * fd + fwd * fd
    return pos, llc, u * 2 * hw * fd, v * 2 * hh * fd


def primary_rays(cam, W, H):
    pos, llc, hor, ver = camera_basis(cam)
    ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    s = 1.0 - (xs.ravel() + 0.5) / W
    t = 1.0 - (ys.ravel() + 0.5) / H
    d = llc[None] + hor[None] * s[:, None] + ver[None] * t[:, None] - pos[None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = d.shape[0]
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = pos
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


def _next_toward(x, sign):
    return np.where(sign > 0, np.nextafter(x, np.float32(np.inf)),
                    np.where(sign < 0, np.nextafter(x, np.float32(-np.inf)), x)).astype(np.float32)


def secondary_rays(tris, rays, t, tri, n_out, seed=4):
    """Cosine-distributed directions about Ng from the hit points of `rays`;
    misses are replaced by other hits (so n_out rays are produced)."""
    rng = np.random.default_rng(seed)
    hit = np.nonzero(tri != 0xFFFFFFFF)[0]
    pick = hit[rng.integers(0, len(hit), n_out)]
    o = rays[pick, 0:3]
    d = rays[pick, 4:7]
    p = (o + d * t[pick, None]).astype(np.float32)
    v = tris[tri[pick]]
    ng = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]).astype(np.float32)
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    # face the side the primary ray came from
    flip = np.sum(ng * d, axis=1) > 0
    ns = np.where(flip[:, None], -ng, ng)
    u1, u2 = rng.random(n_out), rng.random(n_out)
    r = np.sqrt(u1)
    phi = 2 * np.pi * u2
    lx, ly, lz = r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0, 1 - u1))
    a = np.where(np.abs(ns[:, 0:1]) > 0.9, np.array([[0, 1, 0]]), np.array([[1, 0, 0]]))
    tx = np.cross(a, ns)
    tx /= np.linalg.norm(tx, axis=1, keepdims=True)
    ty = np.cross(ns, tx)
    wi = (tx * lx[:, None] + ty * ly[:, None] + ns * lz[:, None]).astype(np.float32)
    perr = (np.abs(p) * GAMMA5).astype(np.float32)
    dd = np.sum(np.abs(ng) * perr, axis=1)
    off = ng * dd[:, None]
    off = np.where((np.sum(wi * ng, axis=1) < 0)[:, None], -off, off).astype(np.float32)
    po = (p + off).astype(np.float32)
    po = _next_toward(po, np.sign(off))
    out = np.zeros((n_out, 8), np.float32)
    out[:, 0:3] = po
    out[:, 3] = 1e-4
    out[:, 4:7] = wi
    out[:, 7] = np.inf
    return out
