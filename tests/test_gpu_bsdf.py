"""Device BxDF evaluate/sample/pdf against the fp64 oracle (<= 1e-5 relative).

The oracle's fp64 BxDFs restate shading/bxdfs/lambert.cpp,
shading/bxdfs/microfacetreflection.cpp, shading/microfacets/ggx.cpp,
shading/fresnel/fresnelconductor.cpp and shading/closures/fresnelspecular.cpp.
Queries are in the local shading frame; sampling decisions that sit within
fp32 rounding of a branch point (F vs u for the dielectric, hemisphere tests)
are counted and must be rare; all other outputs must agree to 1e-5."""
import numpy as np
import pytest

import flgpu
import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-5
FLOOR = 1e-6


def _dirs(rng, n, zmin=0.02):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    small = np.abs(v[:, 2]) < zmin
    v[small, 2] = np.sign(v[small, 2] + 1e-9) * zmin
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v.astype(np.float32)


def make_queries(kind, n, seed, vndf=False):
    rng = np.random.default_rng(seed)
    wo, wi = _dirs(rng, n), _dirs(rng, n)
    qs = []
    for i in range(n):
        q = flgpu.BsdfQuery()
        q.kind = kind
        a = float(rng.uniform(0.05, 0.5))
        q.alpha_x = q.alpha_y = a
        for k in range(4):
            q.eta[k] = float(rng.uniform(0.2, 2.5))
            q.k[k] = float(rng.uniform(0.5, 6.0))
            q.r[k] = float(rng.uniform(0.05, 1.0))
            q.t[k] = float(rng.uniform(0.05, 1.0))
        if kind == flgpu.FL_MAT_DIELECTRIC:
            ior = float(rng.uniform(1.3, 1.8))
            for k in range(4):
                q.eta[k] = ior
        q.wo[:] = wo[i].tolist()
        q.wi[:] = wi[i].tolist()
        q.u[0] = float(rng.random())
        q.u[1] = float(rng.random())
        q.flags = flgpu.FL_RENDER_VNDF_PDF if vndf else 0
        qs.append(q)
    return qs


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), FLOOR)


def _check(ctx, kind, n, seed, vndf=False):
    qs = make_queries(kind, n, seed, vndf)
    g = ctx.bsdf_queries(qs)
    o = oracle.bsdf_batch(qs, 64)
    fe = np.array([list(r.f_eval) for r in g]); ofe = np.array([list(r.f_eval) for r in o])
    pe = np.array([r.pdf_eval for r in g]); ope = np.array([r.pdf_eval for r in o])
    assert _rel(fe, ofe).max() <= RTOL, _rel(fe, ofe).max()
    assert _rel(pe, ope).max() <= RTOL, _rel(pe, ope).max()
    gt = np.array([r.sampled_type for r in g]); ot = np.array([r.sampled_type for r in o])
    same = gt == ot
    frac_branch = 1.0 - same.mean()
    assert frac_branch < 1e-3, frac_branch
    ok = same & (ot != 0)
    fs = np.array([list(r.f_sample) for r in g])[ok]; ofs = np.array([list(r.f_sample) for r in o])[ok]
    ps = np.array([r.pdf_sample for r in g])[ok]; ops = np.array([r.pdf_sample for r in o])[ok]
    ws = np.array([list(r.wi_sample) for r in g])[ok]; ows = np.array([list(r.wi_sample) for r in o])[ok]
    assert np.abs(ws - ows).max() <= RTOL
    assert _rel(ps, ops).max() <= RTOL, _rel(ps, ops).max()
    assert _rel(fs, ofs).max() <= RTOL, _rel(fs, ofs).max()
    return frac_branch


def test_lambert_vs_fp64(ctx):
    _check(ctx, flgpu.FL_MAT_LAMBERT, 20000, 1)


def test_conductor_vs_fp64(ctx):
    _check(ctx, flgpu.FL_MAT_CONDUCTOR, 20000, 2)


def test_conductor_vndf_pdf_vs_fp64(ctx):
    _check(ctx, flgpu.FL_MAT_CONDUCTOR, 20000, 3, vndf=True)


def test_dielectric_vs_fp64(ctx):
    _check(ctx, flgpu.FL_MAT_DIELECTRIC, 20000, 4)
