// fl_develop.hip -- the output path: a pixel's mean XYZ to a display value.
//
//   colour space  xyz_to_rgb / xyz_to_aces2065_1   color/spectrumoperations.cpp:58-76
//   tone map      tone_map                         color/spectrumoperations.cpp:254-286
//   gamma         gamma_correct                    color/spectrumoperations.cpp:215-252
//
// One thread per pixel; the film sums are read where the render left them
// (HBM), divided by spp, and only the developed values are copied to the
// host. Arithmetic follows the reference's expressions in their written
// order (the library builds with -ffp-contract=off).
#include <algorithm>
#include <cstring>

#include "fl_internal.h"

using flx::set_error;

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.f), 1.f); }

// filianore::clamp (maths/mathutils.h:51-55): a NaN passes through
__device__ __forceinline__ float ref_clamp(float v, float lo, float hi) {
    if (v < lo) return lo;
    return v > hi ? hi : v;
}

__device__ __forceinline__ void to_space(uint32_t space, const float x[3], float r[3]) {
    if (space == FL_COLOR_ACES2065_1) {   // xyz_to_aces2065_1, :68-76
        r[0] = 1.0498110175f * x[0] + 0.0000000000f * x[1] - 0.0000974845f * x[2];
        r[1] = -0.4959030231f * x[0] + 1.3733130458f * x[1] + 0.0982400361f * x[2];
        r[2] = 0.0000000000f * x[0] + 0.0000000000f * x[1] + 0.9912520182f * x[2];
    } else {                              // xyz_to_rgb, :58-66
        r[0] = 3.240479f * x[0] - 1.537150f * x[1] - 0.498535f * x[2];
        r[1] = -0.969256f * x[0] + 1.875991f * x[1] + 0.041556f * x[2];
        r[2] = 0.055648f * x[0] - 0.204043f * x[1] + 1.057311f * x[2];
    }
}

__device__ __forceinline__ float tone(uint32_t t, float v) {
    if (t == FL_TONEMAP_ACES_BASIC) {
        const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
        return ref_clamp((v * (a * v + b)) / (v * (c * v + d) + e), 0.f, 1.f);
    }
    if (t == FL_TONEMAP_FILMIC) {
        const float X = fmaxf(0.f, v - 0.004f);
        const float result = (X * (6.2f * X + 0.5f)) / (X * (6.2f * X + 1.7f) + 0.06f);
        return powf(result, 2.2f);
    }
    if (t == FL_TONEMAP_REINHARD) {
        const float L_white = 4.f;
        return (v * (1.f + v / (L_white * L_white))) / (1.f + v);
    }
    return v;
}

__device__ __forceinline__ float gamma_of(uint32_t g, float v) {
    if (g == FL_GAMMA_REC_709) {
        if (v >= 0.018f) return (1.099f * powf(v, 0.45f)) - 0.099f;
        return v * (((1.099f * powf(0.018f, 0.45f)) - 0.099f) / 0.018f);
    }
    if (g == FL_GAMMA_1_8) return powf(v, 1.f / 1.8f);
    if (g == FL_GAMMA_2_2) return powf(v, 1.f / 2.2f);
    if (g == FL_GAMMA_4_0) return powf(v, 1.f / 4.f);
    return v;
}

__device__ __forceinline__ uint8_t to_u8(float v) {
    if (!(v == v)) return 0;   // NaN
    return (uint8_t)(clamp01(v) * 255.f + 0.5f);
}

// sums: n pixels of XYZ sums over spp samples (spp = 1: already means)
__global__ void __launch_bounds__(kBlock) k_develop(const float *__restrict__ sums, uint32_t n, float inv_spp,
                                                    fl_develop_params p, float *__restrict__ rgb,
                                                    uchar4 *__restrict__ rgba8) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float x[3] = {sums[3 * (size_t)i] * inv_spp, sums[3 * (size_t)i + 1] * inv_spp,
                        sums[3 * (size_t)i + 2] * inv_spp};
    float c[3];
    to_space(p.color_space, x, c);
    for (int k = 0; k < 3; ++k) c[k] = gamma_of(p.gamma, tone(p.tone_map, c[k]));
    if (rgb) {
        rgb[3 * (size_t)i] = c[0];
        rgb[3 * (size_t)i + 1] = c[1];
        rgb[3 * (size_t)i + 2] = c[2];
    }
    if (rgba8) rgba8[i] = make_uchar4(to_u8(c[0]), to_u8(c[1]), to_u8(c[2]), 255);
}

int check_params(fl_ctx_t ctx, const fl_develop_params *p) {
    if (!p || p->color_space > FL_COLOR_ACES2065_1 || p->tone_map < FL_TONEMAP_NONE || p->tone_map > FL_TONEMAP_FILMIC ||
        p->gamma < FL_GAMMA_NONE || p->gamma > FL_GAMMA_4_0)
        return set_error(ctx, FL_E_INVALID, "fl_develop: bad parameters");
    return FL_OK;
}

int develop_buffer(fl_ctx_t ctx, size_t bytes) {
    if (ctx->develop_bytes >= bytes) return FL_OK;
    if (ctx->develop_buf) (void)hipFree(ctx->develop_buf);
    ctx->develop_buf = nullptr;
    ctx->develop_bytes = 0;
    if (hipMalloc(&ctx->develop_buf, bytes) != hipSuccess) return set_error(ctx, FL_E_OOM, "fl_develop: staging");
    ctx->develop_bytes = bytes;
    return FL_OK;
}

// develops n pixels from device sums into the staging buffer and copies the
// requested outputs to the host
int develop_device(fl_ctx_t ctx, const float *d_sums, uint32_t n, float inv_spp, const fl_develop_params *p,
                   size_t in_bytes, float *rgb, uint8_t *rgba8) {
    const size_t rgb_bytes = rgb ? (size_t)n * 12 : 0, u8_bytes = rgba8 ? (size_t)n * 4 : 0;
    int rc = develop_buffer(ctx, in_bytes + rgb_bytes + u8_bytes + 256);
    if (rc != FL_OK) return rc;
    char *base = static_cast<char *>(ctx->develop_buf);
    float *d_rgb = rgb ? reinterpret_cast<float *>(base + in_bytes) : nullptr;
    uchar4 *d_u8 = rgba8 ? reinterpret_cast<uchar4 *>(base + in_bytes + rgb_bytes) : nullptr;
    hipStream_t st = ctx->stream;
    if (n) hipLaunchKernelGGL(k_develop, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, d_sums, n, inv_spp, *p,
                              d_rgb, d_u8);
    FL_HIP_CHECK(ctx, hipGetLastError());
    if (rgb) FL_HIP_CHECK(ctx, hipMemcpyAsync(rgb, d_rgb, rgb_bytes, hipMemcpyDeviceToHost, st));
    if (rgba8) FL_HIP_CHECK(ctx, hipMemcpyAsync(rgba8, d_u8, u8_bytes, hipMemcpyDeviceToHost, st));
    FL_HIP_CHECK(ctx, hipStreamSynchronize(st));
    return FL_OK;
}

} // namespace

extern "C" {

int fl_develop_film(fl_ctx_t ctx, uint32_t item, const fl_develop_params *params, float *rgb, uint8_t *rgba8) {
    if (!ctx) return FL_E_INVALID;
    int rc = check_params(ctx, params);
    if (rc != FL_OK) return rc;
    const float *d_sum = nullptr;
    uint32_t W = 0, H = 0, spp = 0;
    rc = flx::last_film(ctx, item, &d_sum, &W, &H, &spp);
    if (rc != FL_OK) return rc;
    (void)hipSetDevice(ctx->device);
    return develop_device(ctx, d_sum, W * H, 1.f / (float)std::max(1u, spp), params, 0, rgb, rgba8);
}

int fl_develop(fl_ctx_t ctx, const float *xyz, uint32_t n, const fl_develop_params *params, float *rgb,
               uint8_t *rgba8) {
    if (!ctx || (!xyz && n)) return set_error(ctx, FL_E_INVALID, "fl_develop: null argument");
    int rc = check_params(ctx, params);
    if (rc != FL_OK) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t in_bytes = ((size_t)n * 12 + 255) / 256 * 256;
    rc = develop_buffer(ctx, in_bytes + (size_t)n * 16 + 256);
    if (rc != FL_OK) return rc;
    FL_HIP_CHECK(ctx, hipMemcpyAsync(ctx->develop_buf, xyz, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
    return develop_device(ctx, static_cast<const float *>(ctx->develop_buf), n, 1.f, params, in_bytes, rgb, rgba8);
}

} // extern "C"
