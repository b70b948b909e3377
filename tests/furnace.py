"""Furnace scenes with closed-form answers for the multi-bounce path loop
(PathIntegrator::Li, integrators/pathintegrator.cpp:12-87).

A closed box whose every triangle is both a Lambert surface of spectral
albedo rho and a one-sided area light of radiance Le facing inward. From any
point inside, every direction sees an emitter of radiance Le, so the direct
light at a vertex is rho * Le in expectation, whichever of light sampling,
BSDF sampling or their MIS (integrator.cpp:25-110) estimates it, and the
cosine-sampled continuation multiplies beta by exactly rho. With the
reference's depth rule (the diffuse break at `bounces == diffuseDepth`
after that vertex's NEE, pathintegrator.cpp:55-65) and emission only at
bounce 0 or after a specular bounce (:23-27):

    L(D) = Le * (1 + rho + rho^2 + ... + rho^(D+1)),

per wavelength: the camera vertex's emission, plus rho^(b+1) Le from the NEE
of vertex b = 0..D. Russian roulette (:77-83) changes the variance, not the
expectation, whatever q it picks.

The glass variant adds a closed lossless dielectric slab (ks = kt = 1,
FresnelSpecularBXDF, fresnelspecular.cpp:8-53). A lossless dielectric in a
uniform radiance field is invisible: the reflected share F and the
transmitted share 1 - F of the uniform radiance sum to it, and the
(eta_i/eta_t)^2 factors of entering and leaving cancel. So without a depth
limit L = Le / (1 - rho) everywhere. Light reaching a wall vertex through the
glass is occluded for NEE (the glass is no emitter,
VisibilityEvaluator::unoccluded, illuminant.cpp:40-50) and is counted instead
by the emission after the specular bounce; a mistake in either half, in the
F split or in the transmission weight moves the mean.

With black walls (rho = 0) every camera sample through the glass returns
exactly Le: beta * etaScale stays 1 inside lossless glass (beta gains
1/eta^2 entering, etaScale eta^2; the reverse leaving), so with a roulette
threshold below 1 roulette never fires and glass pixels have no more
variance than pixels that see the wall directly. A missing or inverted
etaScale (:70-73) sends beta * etaScale to 1/eta^2 or 1/eta^4 inside the
glass and roulette adds variance.
"""
import numpy as np

import flgpu
import oracle

Y_INTEGRAL = 106.856895
LE = 1.0
IOR = 1.5


def box_triangles(lo, hi, inward):
    """The 12 triangles of the box [lo, hi]^3, each face's normal
    (p1 - p0) x (p2 - p0) pointing into the box (inward) or out of it."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    c = (lo + hi) / 2
    tris = []
    for ax in range(3):
        for side in (0, 1):
            u, v = [a for a in range(3) if a != ax]
            q = []
            for (i, j) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                p = np.zeros(3)
                p[ax] = (lo, hi)[side][ax]
                p[u] = (lo, hi)[i][u]
                p[v] = (lo, hi)[j][v]
                q.append(p)
            outward = np.zeros(3)
            outward[ax] = 1.0 if side else -1.0
            want = -outward if inward else outward
            for t in ((q[0], q[1], q[2]), (q[0], q[2], q[3])):
                n = np.cross(t[1] - t[0], t[2] - t[0])
                tris.append(t if np.dot(n, want) > 0 else (t[0], t[2], t[1]))
    out = np.array(tris, np.float32)
    assert np.allclose(out.reshape(-1, 3).mean(0), c)
    return out


def furnace(rho=0.8, glass=False, W=16, H=16, vfov=90.0, spp=16, depth=8):
    """The furnace box [-1, 1]^3 (12 emitting Lambert triangles), optionally
    with the glass slab [-0.5, 0.5]^2 x [-0.7, -0.1]; camera at z = 0.8
    looking down -z. Emission table rows: 0 = Le's spectrum (1 everywhere),
    1 = rho (constant), 2 = 1 (the glass's ks and kt)."""
    walls = box_triangles((-1, -1, -1), (1, 1, 1), inward=True)
    parts = [walls]
    m_wall = flgpu.Material()
    m_wall.kind = flgpu.FL_MAT_LAMBERT
    m_wall.weight = 1.0
    m_wall.spd_r = 2
    mats = [m_wall]
    tri_mat = [0] * 12
    if glass:
        parts.append(box_triangles((-0.5, -0.5, -0.7), (0.5, 0.5, -0.1), inward=False))
        m_g = flgpu.Material()
        m_g.kind = flgpu.FL_MAT_DIELECTRIC
        m_g.ior = IOR
        m_g.spd_r = m_g.spd_t = 3
        mats.append(m_g)
        tri_mat += [1] * 12
    verts = np.concatenate(parts).reshape(-1, 3)
    idx = np.arange(len(verts), dtype=np.uint32).reshape(-1, 3)
    spds = np.stack([np.ones(flgpu.FL_SPD_N), np.full(flgpu.FL_SPD_N, rho), np.ones(flgpu.FL_SPD_N)]).astype(np.float32)
    lights = [(t, 0, LE) for t in range(12)]
    cam = flgpu.Camera()
    cam.position[:] = [0.0, 0.0, 0.8]
    cam.look_at[:] = [0.0, 0.0, -1.0]
    cam.up[:] = [0.0, 1.0, 0.0]
    cam.vfov_deg = vfov
    cam.width, cam.height = W, H
    return flgpu.GenSceneData.from_arrays(verts, idx, np.array(tri_mat, np.uint32), mats, lights, spds, cam,
                                          spp=spp, depth=depth)


def spd_bins(value):
    """The 60-bin spectrum of a constant SPD, as the tables average it."""
    lam = np.arange(360, 831, dtype=np.float32)
    return oracle.from_spd(lam, np.full(lam.shape, value, np.float32)).astype(np.float64)


def parity_xyz(L_bins):
    """XYZ of a 60-bin spectrum: sum_i cmf_i L_i (300 / 60) / CIE_Y_integral
    (to_xyz, spectrumoperations.cpp:58-66), with the bin-averaged CMFs."""
    x, y, z = (np.asarray(c, np.float64) for c in oracle.cie_bins())
    return np.array([np.dot(c, L_bins) for c in (x, y, z)]) * 5.0 / Y_INTEGRAL


def lambert_furnace_bins(rho, D):
    """Per-bin radiance Le * sum_{k=0}^{D+1} rho^k of the Lambert furnace at
    diffuse depth D (the geometric series written out term by term)."""
    r = spd_bins(rho)
    e = spd_bins(1.0) * LE
    total = np.zeros(60)
    term = np.ones(60)
    for _ in range(D + 2):
        total += term
        term = term * r
    return e * total


def expected_parity(rho, D):
    return parity_xyz(lambert_furnace_bins(rho, D))


def expected_glass_parity(rho):
    """Le / (1 - rho) per bin (no depth limit), times the glass's ks = kt,
    which are 1 to float rounding (checked by the caller)."""
    r = spd_bins(rho)
    return parity_xyz(spd_bins(1.0) * LE / (1.0 - r))


def expected_hero(rho, D, cie):
    """Hero mode (lambda_0 uniform in [360, 830], 1-nm tables linearly
    interpolated): each lane's expectation of cmf(lambda) L is the trapezoid
    integral of the 1-nm CMF table over the range divided by 470, and the
    four lanes with weight 470 / (4 Y_INTEGRAL) sum to
    L * trapz(cmf) / Y_INTEGRAL for a spectrally constant L."""
    s = sum(rho ** k for k in range(D + 2)) * LE
    return np.array([np.trapezoid(cie[c].astype(np.float64), dx=1.0) for c in ("FL_CIE_X", "FL_CIE_Y", "FL_CIE_Z")]) \
        * s / Y_INTEGRAL


def image_mean_z(s, q, n, expected, mask=None):
    """z of the image mean of per-pixel means against `expected` (3,), from
    the per-pixel sample variances: s, q are (H, W, 3) sums and sums of
    squares over n samples per pixel. Returns (z (3,), mean (3,), sigma (3,))."""
    s = s.astype(np.float64).reshape(-1, 3)
    q = q.astype(np.float64).reshape(-1, 3)
    if mask is not None:
        s, q = s[mask.ravel()], q[mask.ravel()]
    m = s / n
    var = np.maximum(q / n - m * m, 0.0) * n / (n - 1)
    P = m.shape[0]
    mean = m.mean(0)
    sigma = np.sqrt((var / n).sum(0)) / P
    return (mean - expected) / np.maximum(sigma, 1e-30), mean, sigma


def two_sample_z(a, b):
    """z of the difference of two image means, each (mean, sigma) from image_mean_z."""
    (ma, sa), (mb, sb) = a, b
    return (ma - mb) / np.sqrt(sa * sa + sb * sb)


def glass_pixels(gen):
    """Pixels whose camera ray at the pixel's whole footprint meets the glass
    slab's front face (z = -0.1, |x|, |y| < 0.5) well inside its edges, and
    pixels whose footprint misses the slab entirely: two boolean (H, W) masks."""
    from raysets import camera_basis
    cam = gen.camera
    pos, llc, hor, ver = camera_basis(cam)
    W, H = cam.width, cam.height
    inside = np.zeros((H, W), bool)
    outside = np.zeros((H, W), bool)
    for py in range(H):
        for px in range(W):
            pts = []
            for jx in (0.0, 1.0):
                for jy in (0.0, 1.0):
                    s_ = 1.0 - (px + jx) / W
                    t_ = 1.0 - (py + jy) / H
                    d = llc + hor * s_ + ver * t_ - pos
                    t = (-0.1 - pos[2]) / d[2]
                    pts.append(pos + d * t)
            pts = np.array(pts)
            inside[py, px] = np.all(np.abs(pts[:, :2]) < 0.45)
            # rays diverge from the camera axis, so a footprint wholly beyond
            # |x| or |y| = 0.55 at the front face misses the slab's sides too
            outside[py, px] = any(np.all(pts[:, k] * sg > 0.55) for k in (0, 1) for sg in (1.0, -1.0))
    return inside, outside
