// fl_texture.h -- ImageMapTexture (shading/textures/imagemap.cpp:38-62) on the
// device: a nearest-texel lookup in an 8-bit RGB image at the hit's uv, turned
// into a spectrum by from_reflectance_rgb (color/spectrumoperations.cpp:113-147)
// on the scene's spectral grid.
#pragma once
#include "fl_shade.h"

namespace fl {

struct DevTexture {          // 16 bytes
    uint32_t offset;         // first byte in TexScene::data
    uint32_t width, height;  // 0 x 0: black (an image that did not load)
    uint32_t pad;
};

// A scene's textures and texture coordinates (device memory).
struct TexScene {
    const DevTexture *info;
    const uint8_t *data;     // RGB8 texels of every texture, each 4-byte aligned
    const float2 *tri_uv;    // 3 per triangle (by triangle ID); null: uv = (0, 0)
};

// uv of a hit on triangle `tri` with barycentric weights (u, v) of its second
// and third vertices: Triangle::intersect's area-weighted sum of the vertex
// uvs (shapes/triangle.cpp:66-73); (0, 0) without per-vertex uvs, as the
// Embree path leaves it (bvh/bvh.cpp:69-71).
__device__ __forceinline__ float2 hit_uv(const TexScene &T, uint32_t tri, float u, float v) {
    if (!T.tri_uv) return make_float2(0.f, 0.f);
    const float2 a = T.tri_uv[3ull * tri], b = T.tri_uv[3ull * tri + 1], c = T.tri_uv[3ull * tri + 2];
    const float w = 1.f - u - v;
    return make_float2(a.x * w + b.x * u + c.x * v, a.y * w + b.y * u + c.y * v);
}

// ImageMapTexture::evaluate's texel, rgb in [0, 1]. i and j are clamped to
// the image (the reference's j > height - 1 branch clamps i instead,
// imagemap.cpp:54-55, and then reads past the image).
__device__ __forceinline__ void texel_rgb(const TexScene &T, uint32_t tex, float2 uv, float rgb[3]) {
    const DevTexture D = T.info[tex];
    if (D.width == 0 || D.height == 0) {
        rgb[0] = rgb[1] = rgb[2] = 0.f;
        return;
    }
    int i = (int)(uv.x * (float)D.width);
    int j = (int)((1.f - uv.y) * (float)D.height - 0.001f);
    i = i < 0 ? 0 : (i > (int)D.width - 1 ? (int)D.width - 1 : i);
    j = j < 0 ? 0 : (j > (int)D.height - 1 ? (int)D.height - 1 : j);
    const uint8_t *px = T.data + D.offset + 3ull * ((size_t)i + (size_t)D.width * (size_t)j);
    rgb[0] = (float)(int)px[0] / 255.f;
    rgb[1] = (float)(int)px[1] / 255.f;
    rgb[2] = (float)(int)px[2] / 255.f;
}

// from_reflectance_rgb(r, g, b) at lane k: the white + two-basis combination
// of the 7 reflectance basis curves (white, cyan, magenta, yellow, red, green,
// blue, GS floats apart), scaled by .94 and clamped at 0 at the two table
// entries around the lane's wavelength (as the host forms a material's
// spectrum on the table grid and the material clamps it), then interpolated.
__device__ __forceinline__ float rgb_spectrum_lane(const float *basis, int GS, float r, float g, float bl, const Lanes &ln,
                                                   int k) {
    int b1, b2;
    float c0, c1, c2;
    if (r <= g && r <= bl) {
        c0 = r;
        if (g <= bl) { b1 = 1; c1 = g - r; b2 = 6; c2 = bl - g; }
        else { b1 = 1; c1 = bl - r; b2 = 5; c2 = g - bl; }
    } else if (g <= r && g <= bl) {
        c0 = g;
        if (r <= bl) { b1 = 2; c1 = r - g; b2 = 6; c2 = bl - r; }
        else { b1 = 2; c1 = bl - g; b2 = 4; c2 = r - bl; }
    } else {
        c0 = bl;
        if (r <= g) { b1 = 3; c1 = r - bl; b2 = 5; c2 = g - r; }
        else { b1 = 3; c1 = g - bl; b2 = 4; c2 = r - g; }
    }
    const float *w = basis, *s1 = basis + (size_t)b1 * GS, *s2 = basis + (size_t)b2 * GS;
    float v[2];
    for (int e = 0; e < 2; ++e) {
        const int i = ln.idx[k] + e;
        float x = 0.f + w[i] * c0;
        x = x + s1[i] * c1;
        x = x + s2[i] * c2;
        x = x * 0.94f;
        v[e] = x < 0.f ? 0.f : x;   // spectrum_clamp
    }
    return v[0] + ln.frac[k] * (v[1] - v[0]);
}

} // namespace fl
