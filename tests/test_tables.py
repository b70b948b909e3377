"""Spectral tables of the oracle, pinned independently of its code.

* the 60-bin averages (InitializeSpectrals, color/spectruminits.cpp:6-64;
  average_spectrum_samples, color/spectrumoperations.cpp:28-56) equal the exact
  integral of the piecewise-linear table over each 5 nm bin, computed here
  with numpy on the breakpoints;
* from_reflectance_rgb (spectrumoperations.cpp:113-147) is checked against the
  branch rule written out by hand for specific colours;
* to_xyz of an equal-energy spectrum equals the integral of the CMFs;
* the extracted data file equals the reference headers' float literals.
"""
import os
import re

import numpy as np
import pytest

import flgpu
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src/filianore/color"
CIE_LAM = np.arange(360, 831, dtype=np.float64)


def load_tables():
    text = open(os.path.join(ROOT, "filianore-core_amd", "data", "fl_tables.inc")).read()
    out = {}
    for name, body in re.findall(r"static const float (\w+)\[\d+\] = \{(.*?)\};", text, re.S):
        out[name] = np.array([float.fromhex(t.strip().rstrip("f")) for t in body.split(",") if t.strip()], np.float32)
    return out


T = load_tables()


def exact_bin_average(lam, val, a, b):
    """Average of the piecewise-linear function (flat outside) over [a, b]."""
    pts = np.unique(np.concatenate([[a, b], lam[(lam > a) & (lam < b)]]))
    v = np.interp(pts, lam, val)
    return np.sum(0.5 * (v[1:] + v[:-1]) * np.diff(pts)) / (b - a)


def bins(lam, val):
    edges = np.linspace(400, 700, 61)
    return np.array([exact_bin_average(lam, val, edges[i], edges[i + 1]) for i in range(60)])


def test_cie_bins_match_exact_integrals():
    x, y, z = oracle.cie_bins()
    for got, name in [(x, "FL_CIE_X"), (y, "FL_CIE_Y"), (z, "FL_CIE_Z")]:
        want = bins(CIE_LAM, T[name].astype(np.float64))
        np.testing.assert_allclose(got, want, rtol=5e-6, atol=1e-9)   # fp32 accumulation


def test_rgb_basis_bins_match_exact_integrals():
    lam = T["FL_RGB2SPEC_LAMBDA"].astype(np.float64)
    for i, name in enumerate(["WHITE", "CYAN", "MAGENTA", "YELLOW", "RED", "GREEN", "BLUE"]):
        want = bins(lam, T["FL_RGBREFL_" + name].astype(np.float64))
        np.testing.assert_allclose(oracle.refl_basis(i), want, rtol=1e-5, atol=1e-7)


def test_metal_bins_match_exact_integrals():
    for m, name in enumerate(["GOLD", "COPPER", "SILVER", "ALUMINIUM"]):
        lam = T["FL_METAL_%s_LAMBDA" % name].astype(np.float64)
        eta, k = oracle.metal_bins(m)
        np.testing.assert_allclose(eta, bins(lam, T["FL_METAL_%s_ETA" % name]), rtol=1e-5)
        np.testing.assert_allclose(k, bins(lam, T["FL_METAL_%s_K" % name]), rtol=1e-5)


def test_from_reflectance_rgb_branch_rule():
    W, Cy, M, Ye, R, G, B = (oracle.refl_basis(i).astype(np.float64) for i in range(7))
    cases = {
        (0.2, 0.5, 0.8): 0.2 * W + 0.3 * Cy + 0.3 * B,      # r <= g <= b
        (0.2, 0.8, 0.5): 0.2 * W + 0.3 * Cy + 0.3 * G,      # r <= b < g
        (0.5, 0.2, 0.8): 0.2 * W + 0.3 * M + 0.3 * B,       # g <= r <= b
        (0.8, 0.2, 0.5): 0.2 * W + 0.3 * M + 0.3 * R,       # g <= b < r
        (0.5, 0.8, 0.2): 0.2 * W + 0.3 * Ye + 0.3 * G,      # b lowest, r <= g
        (0.8, 0.5, 0.2): 0.2 * W + 0.3 * Ye + 0.3 * R,      # b lowest, g < r
        (1.0, 1.0, 1.0): 1.0 * W,
    }
    for rgb, spec in cases.items():
        want = np.maximum(spec * 0.94, 0)
        np.testing.assert_allclose(oracle.from_reflectance_rgb(rgb), want, rtol=1e-5, atol=1e-6)


def test_to_xyz_equal_energy_is_cmf_integral():
    xyz = oracle.to_xyz(np.ones(60, np.float32))
    sel = (CIE_LAM >= 400) & (CIE_LAM <= 700)
    want = [np.trapezoid(T[n][sel].astype(np.float64), CIE_LAM[sel]) / 106.856895 for n in
            ("FL_CIE_X", "FL_CIE_Y", "FL_CIE_Z")]
    np.testing.assert_allclose(xyz, want, rtol=1e-5)
    assert abs(xyz[1] - 1.0) < 0.01      # 400..700 nm holds almost all of y-bar


def test_blackbody_peak_follows_wien():
    spd = (flgpu.C.c_float * flgpu.FL_SPD_N)()
    flgpu.lib().fl_gen_blackbody(6500.0, spd)
    a = np.frombuffer(spd, np.float32)
    assert abs(a.max() - 1.0) < 1e-6
    lam_peak = 360 + int(np.argmax(a))
    assert abs(lam_peak - 2.897771955e6 / 6500.0) <= 1.0


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_extracted_tables_equal_reference_literals():
    num = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")

    def ref_array(fname, name):
        text = re.sub(r"//[^\n]*", "", open(os.path.join(REF, fname)).read())
        body = re.search(r"const float %s\s*\[[^\]]*\]\s*=\s*\{(.*?)\};" % name, text, re.S).group(1)
        return np.array([float(t) for t in num.findall(body)], np.float32)

    assert np.array_equal(T["FL_CIE_Y"], ref_array("cie.h", "CIE_Y"))
    assert np.array_equal(T["FL_RGBREFL_BLUE"], ref_array("cie.h", "RGBRefl2SpectBlue"))
    assert np.array_equal(T["FL_METAL_COPPER_K"], ref_array("metallurgy.h", "Copper_K"))
    assert np.array_equal(T["FL_METAL_LITHIUM_LAMBDA"], ref_array("metallurgy.h", "Lithium_Lambda"))
