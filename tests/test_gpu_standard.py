"""StandardSurfaceMaterial on the GPU (shading/materials/standardsurface.cpp:72-159):

* fl_bsdf_query_standard (BSDF::evaluate / pdf without specular lobes and
  BSDF::sample over all lobes, bsdf.cpp:12-151) against the fp64 oracle for
  each lobe alone and for stacks: f and pdf within 1e-5 relative (values
  under 1e-6 compared absolutely); sampled lobe types agree except where u
  sits within fp32 rounding of a branch point (< 0.1% of queries), and
  agreeing samples give the same direction to 1e-5. A sample's pdf and f
  must match, within 1e-5, the fp64 oracle's sample or the oracle's pdf and
  f at the device's own sampled direction (times (n - 1) / n for the pdf
  when the stack holds the specular lobe, which BSDF::sample counts and the
  non-specular pdf does not). Near grazing, a lobe's pdf moves by more than
  1e-5 under the fp32 rounding of the direction alone: the sampled lobe's
  pdf comes from its sampled half vector, the others' from the direction,
  and only one of the two comparisons sees the device's own rounding;
* a Cornell box whose walls and blocks are standard surfaces (every lobe
  kind, thin-walled translucency with a specular dielectric lobe) against
  the oracle's render at equal spp, per-pixel z-test as in
  test_gpu_render.py, in parity mode; hero mode's image mean within a few
  percent of parity mode's."""
import numpy as np
import pytest

import flgpu
import oracle
from stdsurf import MATERIALS, has_specular_lobe, material, n_lobes, queries, random_dirs, surface
from test_gpu_render import z_stats

pytestmark = pytest.mark.gpu

RTOL = 1e-5
FLOOR = 1e-6


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), FLOOR)


def _check(ctx, m, n, seed, vndf):
    rng = np.random.default_rng(seed)
    q = queries(random_dirs(rng, n), random_dirs(rng, n), u=rng.random((n, 2)), vndf=vndf)
    g = ctx.bsdf_queries_standard(m, q)
    o = oracle.bsdf_standard(m, q, 64)
    same = g["sampled_type"] == o["sampled_type"]
    assert 1.0 - same.mean() < 1e-3, 1.0 - same.mean()
    ok = same & (o["sampled_type"] != 0)
    assert ok.mean() > 0.3
    assert np.abs(g["wi_sample"][ok] - o["wi_sample"][ok]).max() <= RTOL
    spec = ok & ((o["sampled_type"] & flgpu.BSDF_SPECULAR) != 0)
    ns = ok & ~spec
    q2 = q[ns].copy()
    q2["wi"] = g["wi_sample"][ns]
    o2 = oracle.bsdf_standard(m, q2, 64)
    n = n_lobes(m)
    ratio = (n - 1) / n if has_specular_lobe(m) else 1.0
    # each sample against the oracle's own sample and against the oracle's
    # pdf / f at the device's direction; the better of the two must hold
    dp = _rel(g["pdf_sample"][ok], o["pdf_sample"][ok])
    df = _rel(g["f_sample"][ok], o["f_sample"][ok]).max(-1)
    rp = np.full(dp.shape, np.inf)
    rf = np.full(dp.shape, np.inf)
    sel = ns[ok]
    rp[sel] = _rel(g["pdf_sample"][ns], o2["pdf_eval"] * ratio)
    rf[sel] = _rel(g["f_sample"][ns], o2["f_eval"]).max(-1)
    ep, ef = np.minimum(dp, rp), np.minimum(df, rf)
    print("f %.1e pdf %.1e | sample pdf %.1e f %.1e (direct %.1e %.1e)" % (
        _rel(g["f_eval"], o["f_eval"]).max(), _rel(g["pdf_eval"], o["pdf_eval"]).max(),
        ep.max(), ef.max(), dp.max(), df.max()))
    assert _rel(g["f_eval"], o["f_eval"]).max() <= RTOL, _rel(g["f_eval"], o["f_eval"]).max()
    assert _rel(g["pdf_eval"], o["pdf_eval"]).max() <= RTOL, _rel(g["pdf_eval"], o["pdf_eval"]).max()
    assert ep.max() <= RTOL and ef.max() <= RTOL, (ep.max(), ef.max())
    # the oracle's own sample, within the conditioning of the direction
    assert dp.max() <= 1e-3 and df.max() <= 1e-3
    return o


@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_standard_bsdf_vs_fp64(ctx, name):
    o = _check(ctx, material(name), 20000, sum(map(ord, name)), vndf=False)
    assert np.any(o["f_eval"] > 0) or name == "fspec"


@pytest.mark.parametrize("name", ["conductor", "fblend_metallic", "oren+conductor+thinfilm+coat"])
def test_standard_bsdf_vndf_pdf_vs_fp64(ctx, name):
    _check(ctx, material(name), 20000, 7, vndf=True)


def test_bad_fresnel_type_is_rejected(ctx):
    m = surface(ks_weight=1.0, fresnel_type=7)
    with pytest.raises(flgpu.FlError):
        ctx.bsdf_queries_standard(m, queries(np.array([[0, 0, 1.0]])))


def standard_box(width=32, height=32):
    """The Cornell box with StandardSurface materials: white walls with a
    blended diffuse/specular base, thin film, sheen and coat; one coloured wall
    rough Oren-Nayar under a wavelength-dependent metal, thin film and coat;
    the other a Schlick-metallic blend; the blocks thin-walled translucent
    (diffuse transmission, specular dielectric, coat)."""
    gen = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=width, height=height)
    names = ["fblend+thinfilm+sheen+coat", "oren+conductor+thinfilm+coat", "fblend_metallic", "difftrans+fspec+coat"]
    colours = [(0.75, 0.72, 0.68), (0.65, 0.12, 0.10), (0.12, 0.45, 0.15), (0.7, 0.7, 0.6)]
    ss = []
    for i, (name, rgb) in enumerate(zip(names, colours)):
        s = material(name)
        s.kd[:] = rgb
        s.ks[:] = (0.9, 0.85, 0.8)
        s.kt[:] = (0.8, 0.9, 0.95)
        s.sheen_color[:] = (0.9, 0.6, 0.4)
        s.coat[:] = (0.95, 0.95, 1.0)
        s.metal = i * 2 % 9
        ss.append(s)
        gen.materials[i].kind = flgpu.FL_MAT_STANDARD
        gen.materials[i].standard = i
    gen.standard_surfaces = (flgpu.StandardSurface * len(ss))(*ss)
    gen.n_standard_surfaces = len(ss)
    return gen


def test_standard_box_consistent_with_oracle(ctx):
    gen = standard_box()
    spp = 1024
    with ctx.scene(gen) as sc:
        p = flgpu.default_params(spp=spp, depth=gen.depth, mode=flgpu.FL_SPECTRAL_PARITY, seed=51)
        gs, gq, st = sc.render(p)
    assert st.nan_samples == 0
    os_, oq, ost = oracle.Scene(gen).render(spp, gen.depth, seed=52, threads=16)
    frac, allowed, mean_z, bad = z_stats(gs, gq, os_, oq, spp)
    print("standard box: frac %.4f allowed %.4f mean z %.2f cpu %.1fs" % (frac, allowed, mean_z, ost.seconds))
    assert bad == 0 and frac <= allowed and abs(mean_z) < 3
    # the standard surfaces change the image: it is not the Lambert box
    plain = flgpu.GenSceneData(flgpu.FL_GEN_CORNELL, width=32, height=32)
    ps, _, _ = oracle.Scene(plain).render(64, plain.depth, seed=53, threads=16)
    assert abs(os_.mean() / spp - ps.mean() / 64) > 0.05 * (ps.mean() / 64)


def test_standard_box_hero_close_to_parity(ctx):
    gen = standard_box(48, 48)
    with ctx.scene(gen) as sc:
        a, _, sa = sc.render(flgpu.default_params(spp=256, depth=gen.depth, mode=flgpu.FL_SPECTRAL_HERO, seed=3))
        b, _, sb = sc.render(flgpu.default_params(spp=256, depth=gen.depth, mode=flgpu.FL_SPECTRAL_PARITY, seed=4))
    assert sa.nan_samples == 0 and sb.nan_samples == 0
    ma, mb = a.reshape(-1, 3).mean(0), b.reshape(-1, 3).mean(0)
    assert np.all(np.abs(ma - mb) / mb < 0.05), (ma, mb)
