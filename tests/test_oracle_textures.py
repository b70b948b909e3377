"""Image textures (ImageMapTexture, shading/textures/imagemap.cpp:38-62) in the
CPU oracle, pinned by a closed form: a floor lit straight from above by a
directional light of intensity I has L = rho(uv)/pi * I, where rho(uv) is
from_reflectance_rgb of the texel that the reference's lookup picks,
i = int(u W), j = int((1 - v) H - 0.001) (rows top to bottom). Each pixel
averages that over a 16 x 16 grid of floor points inside it, as the jittered
camera samples do; uv comes from per-vertex coordinates interpolated as
Triangle::intersect does, or is (0, 0) on every hit without them (the
Embree path), which picks the bottom-left texel everywhere. The same holds
with the texture as a StandardSurface's kd (Oren-Nayar of roughness 0 is
Lambert)."""
import numpy as np
import pytest

import oracle
from scenes import checker_texture, floor_points, texel_of, textured_floor
from test_oracle_lights import assert_close, lit_xyz, render_oracle, spectra


def want_image(gen, texture, intensity, with_uvs):
    P = floor_points(gen.camera, sub=16)
    if with_uvs:
        i, j = texel_of(texture, (P[..., 0] + 2.0) / 4.0, (P[..., 2] + 2.0) / 4.0)
    else:
        i, j = texel_of(texture, np.zeros(P.shape[:-1]), np.zeros(P.shape[:-1]))
    h, w = texture.shape[:2]
    spd = spectra()[0]
    table = np.array([[lit_xyz(spd, texture[jj, ii] / 255.0) for ii in range(w)] for jj in range(h)])
    return table[j, i].mean(-2) * intensity, (i, j)


@pytest.mark.parametrize("standard", [False, True])
def test_textured_floor_closed_form(standard):
    tex = checker_texture()
    gen = textured_floor(tex, spectra(), standard=standard)
    img, sig = render_oracle(gen, with_sigma=True)
    want, (i, j) = want_image(gen, tex, 2.0, True)
    assert_close(img, sig, want)
    assert len(np.unique(i)) == 5 and len(np.unique(j)) == 4   # the view crosses every texel


def test_texture_without_uvs_is_the_bottom_left_texel():
    tex = checker_texture(seed=1)
    gen = textured_floor(tex, spectra(), with_uvs=False)
    img, sig = render_oracle(gen, with_sigma=True)
    want, _ = want_image(gen, tex, 2.0, False)
    assert_close(img, sig, want)
    assert np.allclose(want, want[0, 0])


def test_missing_image_is_black():
    tex = np.zeros((0, 0, 3), np.uint8)
    gen = textured_floor(tex, spectra())
    img = render_oracle(gen)
    assert np.all(img == 0)
