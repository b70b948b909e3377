// fl_math.h -- small vector/scalar helpers shared by host and device code.
//
// Restates the reference's geometry helpers that the render path relies on:
//   dot/cross/normalize/face_forward/reflect/refract  maths/geometry.h:203-328
//   coordinate_system                                 maths/geometry.h:246-256
//   next_float_up/down, gamma(n)                      maths/mathutils.h:40-48,97-124
//   offset_point_origin                               core/interaction.h:11-31
// All HIP sources are compiled with -ffp-contract=off so that every expression
// here rounds exactly as written, on the GPU and on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FL_HD __host__ __device__ __forceinline__

namespace fl {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kPiOver2 = 1.57079632679489661923f;
constexpr float kPiOver4 = 0.78539816339744830961f;
constexpr float kShadowEps = 0.0001f;          // SHADOW_EPSILON, mathutils.h:20
constexpr float kInf = __builtin_huge_valf();

struct v3 {
    float x, y, z;
    FL_HD v3() : x(0), y(0), z(0) {}
    FL_HD v3(float a, float b, float c) : x(a), y(b), z(c) {}
    FL_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

FL_HD v3 operator+(v3 a, v3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
FL_HD v3 operator-(v3 a, v3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
FL_HD v3 operator-(v3 a) { return v3(-a.x, -a.y, -a.z); }
FL_HD v3 operator*(v3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
FL_HD v3 operator*(float s, v3 a) { return v3(a.x * s, a.y * s, a.z * s); }
FL_HD v3 operator/(v3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
FL_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FL_HD float absdot(v3 a, v3 b) { return fabsf(dot(a, b)); }
FL_HD v3 vabs(v3 a) { return v3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
FL_HD v3 cross(v3 a, v3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
FL_HD float length(v3 a) { return sqrtf(dot(a, a)); }
FL_HD float length2(v3 a) { return dot(a, a); }
FL_HD v3 normalize(v3 a) { return a / length(a); }
FL_HD v3 face_forward(v3 v, v3 w) { return dot(v, w) < 0.f ? -v : v; }
FL_HD float fmaxf3(float a, float b, float c) { return fmaxf(a, fmaxf(b, c)); }
FL_HD float fminf3(float a, float b, float c) { return fminf(a, fminf(b, c)); }
FL_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// reflect(a, b) = b*2*dot(a,b) - a          (geometry.h:309-312)
FL_HD v3 reflect(v3 a, v3 b) { return b * 2.f * dot(a, b) - a; }

// refract(wi, n, eta, &wt)                    (geometry.h:314-328)
FL_HD bool refract(v3 wi, v3 n, float eta, v3 *wt) {
    float cosI = dot(wi, n);
    float sinI2 = fmaxf(0.f, 1.f - cosI * cosI);
    float sinT2 = eta * eta * sinI2;
    if (sinT2 >= 1.f) return false;
    float cosT = sqrtf(1.f - sinT2);
    *wt = wi * -eta + n * (eta * cosI - cosT);
    return true;
}

// coordinate_system(v1, &v2, &v3)             (geometry.h:246-256)
FL_HD void coordinate_system(v3 v1, v3 *v2, v3 *v3o) {
    if (fabsf(v1.x) > fabsf(v1.y)) {
        float inv = 1.f / sqrtf(v1.x * v1.x + v1.z * v1.z);
        *v2 = v3(v1.z * inv, 0.f, -v1.x * inv);
    } else {
        float inv = 1.f / sqrtf(v1.y * v1.y + v1.z * v1.z);
        *v2 = v3(0.f, v1.z * inv, -v1.y * inv);
    }
    *v3o = cross(v1, *v2);
}

FL_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
FL_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// next_float_up / next_float_down             (mathutils.h:97-124)
FL_HD float next_up(float v) {
    if (v == kInf) return v;
    if (v == -0.f) v = 0.f;
    uint32_t ui = f2u(v);
    ui = (v >= 0.f) ? ui + 1u : ui - 1u;
    return u2f(ui);
}
FL_HD float next_down(float v) {
    if (v == -kInf) return v;
    if (v == 0.f) v = -0.f;
    uint32_t ui = f2u(v);
    ui = (v > 0.f) ? ui - 1u : ui + 1u;
    return u2f(ui);
}

// gamma<float>(5): pError bound of the hit point (repair of bvh.cpp:78, which
// evaluates gamma<int>(5) == 0; triangle.cpp:127 uses gamma<float>(5)).
constexpr float kMachEps = 5.9604644775390625e-08f; // FLT_EPSILON/2
FL_HD float gamma_n(int n) { return (n * kMachEps) / (1.f - n * kMachEps); }

// offset_point_origin(p, pError, n, w)          (interaction.h:11-31)
FL_HD v3 offset_point_origin(v3 p, v3 perr, v3 n, v3 w) {
    float d = dot(vabs(n), perr);
    v3 off = n * d;
    if (dot(w, n) < 0.f) off = -off;
    v3 po = p + off;
    po.x = off.x > 0.f ? next_up(po.x) : (off.x < 0.f ? next_down(po.x) : po.x);
    po.y = off.y > 0.f ? next_up(po.y) : (off.y < 0.f ? next_down(po.y) : po.y);
    po.z = off.z > 0.f ? next_up(po.z) : (off.z < 0.f ? next_down(po.z) : po.z);
    return po;
}

// power_heuristic(1, f, 1, g)                  (sampling/sampling.cpp:18-23)
FL_HD float power_heuristic(float f, float g) { return (f * f) / (f * f + g * g); }

} // namespace fl
