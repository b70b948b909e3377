// fl_bvh.h -- device layout of the 8-wide quantized BVH and triangle records,
// plus the watertight ray/triangle test used for every closest/any-hit query.
//
// Replaces Embree's role behind BVH::intersect (bvh/bvh.cpp:49-84).  The
// watertight test is Woop, Benthin & Wald, "Watertight Ray/Triangle
// Intersection", JCGT 2(1) 2013; it is written here with a fixed operation
// order and no contraction so that the CPU oracle, which restates it, returns
// the same bits.  Tie rule: at equal t the smaller triangle ID wins.
#pragma once
#include "fl_math.h"

namespace fl {

// 128-byte node: one cache line.  Child boxes are 16-bit quantized against a
// power-of-two grid anchored at `org`, which is itself a multiple of the grid
// step, so every dequantized plane (org + q*2^e) is exact in fp32.
struct alignas(128) Node8 {
    float org[3];
    uint8_t ex, ey, ez;      // biased exponents: step = 2^(e-127)
    uint8_t imask;           // bit s: slot s holds an internal child
    uint32_t child_base;     // index of the first internal child node
    uint32_t tri_base;       // index of the first triangle record of leaf slots
    uint8_t meta[8];         // internal: rank among internal children;
                             // leaf: (offset & 31) | (count << 5), count 1..3; empty: 0
    uint16_t qlo[3][8];      // per axis, per slot
    uint16_t qhi[3][8];
};
static_assert(sizeof(Node8) == 128, "node must be one 128-byte line");

// 48-byte triangle record in leaf order.
struct alignas(16) TriRec {
    float v0[3];
    uint32_t id;             // original (flattened) triangle ID
    float v1[3];
    uint32_t material;       // material index, 0xffff = none (pass-through)
    float v2[3];
    uint32_t light;          // light index + 1, 0 = not emissive
};
static_assert(sizeof(TriRec) == 48, "tri record is 48 bytes");

constexpr uint32_t kNoMaterial = 0xffffu;
constexpr uint32_t kNoHit = 0xffffffffu;

// Per-ray precomputation of the watertight test.
struct RayPre {
    uint32_t k;              // kx | ky << 2 | kz << 4 (one register per ray)
    float Sx, Sy, Sz;
};

FL_HD RayPre ray_pre(v3 d) {
    RayPre r;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    float dkz = d[kz];
    if (dkz < 0.f) { int t = kx; kx = ky; ky = t; }
    r.k = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4);
    r.Sx = d[kx] / dkz;
    r.Sy = d[ky] / dkz;
    r.Sz = 1.0f / dkz;
    return r;
}

FL_HD float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// Returns true with (t, u, v) when the ray hits (v0, v1, v2) with tmin < t <= tmax.
// u, v are the barycentric weights of v1 and v2 (Embree convention).
FL_HD bool tri_intersect(const RayPre &rp, v3 o, v3 p0, v3 p1, v3 p2, float tmin, float tmax,
                         float *t_out, float *u_out, float *v_out) {
    v3 A = p0 - o, B = p1 - o, C = p2 - o;
    const int kx = (int)(rp.k & 3u), ky = (int)((rp.k >> 2) & 3u), kz = (int)(rp.k >> 4);
    float Akx = comp(A, kx), Aky = comp(A, ky), Akz = comp(A, kz);
    float Bkx = comp(B, kx), Bky = comp(B, ky), Bkz = comp(B, kz);
    float Ckx = comp(C, kx), Cky = comp(C, ky), Ckz = comp(C, kz);
    float Ax = Akx - rp.Sx * Akz, Ay = Aky - rp.Sy * Akz;
    float Bx = Bkx - rp.Sx * Bkz, By = Bky - rp.Sy * Bkz;
    float Cx = Ckx - rp.Sx * Ckz, Cy = Cky - rp.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.f || V == 0.f || W == 0.f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    float det = U + V + W;
    if (det == 0.f) return false;
    float Az = rp.Sz * Akz, Bz = rp.Sz * Bkz, Cz = rp.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float inv = 1.0f / det;
    float t = T * inv;
    if (!(t > tmin && t <= tmax)) return false;
    *t_out = t;
    *u_out = V * inv;
    *v_out = W * inv;
    return true;
}

// Closest-hit acceptance with the tie rule (smaller ID at equal t).
FL_HD bool closer(float t, uint32_t id, float best_t, uint32_t best_id) {
    return t < best_t || (t == best_t && id < best_id);
}

} // namespace fl
