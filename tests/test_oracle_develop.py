"""The oracle's output path (xyz_to_rgb / xyz_to_aces2065_1, tone_map,
gamma_correct; color/spectrumoperations.cpp:58-76,215-286) pinned by what
the colour science and the formulas fix: white points map to (1, 1, 1), the
sRGB matrix inverts RGBToXYZ (:78-86), known values of each tone curve,
the continuity of the Rec. 709 curve at its 0.018 break, and the 8-bit
rounding rule."""
import numpy as np
import pytest

import flgpu
import oracle

NONE_T, NONE_G = flgpu.FL_TONEMAP_NONE, flgpu.FL_GAMMA_NONE


def dev(xyz, space=flgpu.FL_COLOR_SRGB, tm=NONE_T, gm=NONE_G):
    return oracle.develop(np.asarray(xyz, np.float32).reshape(-1, 3), space, tm, gm)


def test_white_points_map_to_unit_rgb():
    # D65 (sRGB white) and the ACES AP0 white point (x, y) = (0.32168, 0.33767), Y = 1
    d65 = [0.95047, 1.0, 1.08883]
    x, y = 0.32168, 0.33767
    ap0 = [x / y, 1.0, (1 - x - y) / y]
    rgb, _ = dev(d65)
    np.testing.assert_allclose(rgb[0], [1, 1, 1], atol=2e-3)
    aces, _ = dev(ap0, flgpu.FL_COLOR_ACES2065_1)
    np.testing.assert_allclose(aces[0], [1, 1, 1], atol=2e-4)


def test_srgb_matrix_inverts_rgb_to_xyz():
    # RGBToXYZ of the same file (spectrumoperations.cpp:78-86): the inverse map
    M = np.array([[0.412453, 0.357580, 0.180423], [0.212671, 0.715160, 0.072169], [0.019334, 0.119193, 0.950227]])
    rgb = np.random.default_rng(1).uniform(0, 1, (200, 3))
    out, _ = dev((rgb @ M.T).astype(np.float32))
    np.testing.assert_allclose(out, rgb, atol=2e-5)


def test_tone_curves():
    v = np.array([0.0, 1.0, 4.0, 1e6], np.float32)
    xyz = np.zeros((4, 3), np.float32)
    # put v on the sRGB red channel only: XYZ = v * (first column of RGBToXYZ) gives rgb ~ (v, 0, 0)
    M = np.array([[0.412453, 0.357580, 0.180423], [0.212671, 0.715160, 0.072169], [0.019334, 0.119193, 0.950227]])
    xyz = (v[:, None] * M[:, 0][None, :]).astype(np.float32)
    lin, _ = dev(xyz)
    r = lin[:, 0]
    reinhard, _ = dev(xyz, tm=flgpu.FL_TONEMAP_REINHARD)
    np.testing.assert_allclose(reinhard[:, 0], r * (1 + r / 16) / (1 + r), rtol=1e-6)
    assert abs(reinhard[2, 0] - 1.0) < 1e-4                     # L_white = 4 maps to 1
    aces, _ = dev(xyz, tm=flgpu.FL_TONEMAP_ACES_BASIC)
    assert aces[0, 0] == 0.0 and aces[3, 0] == 1.0              # clamped at 1
    assert abs(aces[1, 0] - 2.54 / 3.16) < 1e-4                 # (a + b) / (c + d + e) at 1
    filmic, _ = dev(xyz, tm=flgpu.FL_TONEMAP_FILMIC)
    assert filmic[0, 0] == 0.0 and abs(filmic[3, 0] - 1.0) < 1e-5
    X = 1.0 - 0.004
    assert abs(filmic[1, 0] - ((X * (6.2 * X + 0.5)) / (X * (6.2 * X + 1.7) + 0.06)) ** 2.2) < 1e-4


def test_gamma_curves():
    g = np.array([0.0179, 0.018, 0.25, 1.0], np.float32)
    M = np.array([[0.412453, 0.357580, 0.180423], [0.212671, 0.715160, 0.072169], [0.019334, 0.119193, 0.950227]])
    xyz = (g[:, None] * M[:, 0][None, :]).astype(np.float32)
    lin, _ = dev(xyz)
    r = lin[:, 0].astype(np.float64)
    rec, _ = dev(xyz, gm=flgpu.FL_GAMMA_REC_709)
    # the power branch at and above 0.018, the linear one below; they meet at 0.018
    knee = 1.099 * 0.018 ** 0.45 - 0.099
    np.testing.assert_allclose(rec[1:, 0], 1.099 * r[1:] ** 0.45 - 0.099, rtol=1e-5)
    np.testing.assert_allclose(rec[0, 0], r[0] * knee / 0.018, rtol=1e-5)
    assert abs(rec[1, 0] - rec[0, 0]) < 1e-3 and abs(rec[3, 0] - 1.0) < 1e-3
    for gm, e in ((flgpu.FL_GAMMA_1_8, 1 / 1.8), (flgpu.FL_GAMMA_2_2, 1 / 2.2), (flgpu.FL_GAMMA_4_0, 0.25)):
        out, _ = dev(xyz, gm=gm)
        np.testing.assert_allclose(out[:, 0], r ** e, rtol=2e-6)


def test_rgba8_rounding():
    M = np.array([[0.412453, 0.357580, 0.180423], [0.212671, 0.715160, 0.072169], [0.019334, 0.119193, 0.950227]])
    Minv = np.linalg.inv(M)
    rgb = np.array([[0.5, -1.0, 2.0], [0.0, 1.0, 0.25]])
    out, u8 = dev((rgb @ M.T).astype(np.float32))
    want = np.where(np.isnan(out), 0, np.floor(np.clip(out, 0, 1) * 255 + 0.5)).astype(np.uint8)
    assert np.array_equal(u8[:, :3], want) and np.all(u8[:, 3] == 255)
    assert u8[0, 1] == 0 and u8[0, 2] == 255
    # a negative value through a fractional power is NaN (powf), stored as 0
    nan_in = (np.array([[-0.5, 0.2, 0.2]]) @ M.T).astype(np.float32)
    out, u8 = dev(nan_in, gm=flgpu.FL_GAMMA_2_2)
    assert np.isnan(out[0, 0]) and u8[0, 0] == 0
