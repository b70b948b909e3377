// fl_trace.h -- per-lane traversal of the 8-wide quantized BVH.
//
// One ray per lane (wave64).  The traversal keeps a "node group" (a node's
// internal children still to visit, as a hit mask over octant-ordered bit
// positions) and a "triangle group" (hit leaf triangles of the last node), in
// the style of compressed wide BVH traversal (Ylitie, Karras, Laine, HPG
// 2017), with a short stack of node groups in LDS that spills to a per-lane
// global area when it is full.  Box tests dequantize planes exactly
// (org + q*step is exact by construction) and widen the slab interval by a
// relative 2^-20, so no box containing a hit is ever culled: the closest hit
// (with the smaller-ID tie rule) is then the same as a brute-force search.
#pragma once
#include "fl_bvh.h"

namespace fl {

struct TraceScene {
    const Node8 *__restrict__ nodes;
    const TriRec *__restrict__ tris;
};

struct TraceHit {
    float t, u, v;
    uint32_t id;   // original triangle ID or kNoHit
    uint32_t rec;  // triangle record index (leaf order)
};

struct LaneStack {
    uint2 *lds;        // lane's first entry; entry i at lds[i * lds_stride]
    int lds_stride;
    int cap;           // entries in LDS
    uint2 *spill;      // lane's first spill entry; entry i at spill[i * spill_stride]
    size_t spill_stride;
};

__device__ __forceinline__ float safe_inv(float d) {
    const float eps = 1e-20f;
    float dd = fabsf(d) > eps ? d : copysignf(eps, d);
    return 1.0f / dd;
}

template <bool ANY, bool SKIP_LIGHTS, bool COUNT>
__device__ __forceinline__ bool traverse(const TraceScene &sc, v3 o, v3 d, float tmin, float tmax,
                                         const LaneStack &ls, TraceHit &hit, uint32_t &cnt_nodes,
                                         uint32_t &cnt_tris, uint32_t &cnt_spill) {
    const float ix = safe_inv(d.x), iy = safe_inv(d.y), iz = safe_inv(d.z);
    const uint32_t oct = (ix < 0.f ? 1u : 0u) | (iy < 0.f ? 2u : 0u) | (iz < 0.f ? 4u : 0u);
    const RayPre rp = ray_pre(d);
    const float kLo = 1.0f - 0x1p-20f, kHi = 1.0f + 0x1p-20f;

    float best_t = tmax;
    uint32_t best_id = kNoHit, best_rec = 0;
    float bu = 0.f, bv = 0.f;

    uint32_t g_base = 0, g_imask = 1u, g_hits = 1u << oct;  // group holding the root
    uint32_t t_base = 0, t_hits = 0;
    int sp = 0;
    bool done = false;

    while (true) {
        if (g_hits) {
            uint32_t pos = __ffs(g_hits) - 1;
            g_hits &= g_hits - 1u;
            uint32_t slot = pos ^ oct;
            uint32_t nidx = g_base + __popc(g_imask & ((1u << slot) - 1u));
            if (g_hits) {
                uint2 e = make_uint2(g_base, g_imask | (g_hits << 8));
                if (sp < ls.cap) ls.lds[sp * ls.lds_stride] = e;
                else { ls.spill[(size_t)(sp - ls.cap) * ls.spill_stride] = e; if (COUNT) ++cnt_spill; }
                ++sp;
            }
            const uint4 *np = reinterpret_cast<const uint4 *>(sc.nodes + nidx);
            uint4 w0 = np[0], w1 = np[1];
            uint4 ql[3] = {np[2], np[3], np[4]};
            uint4 qh[3] = {np[5], np[6], np[7]};
            if (COUNT) ++cnt_nodes;
            float org[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
            float step[3] = {__uint_as_float((w0.w & 0xffu) << 23), __uint_as_float(((w0.w >> 8) & 0xffu) << 23),
                             __uint_as_float(((w0.w >> 16) & 0xffu) << 23)};
            uint32_t imask = w0.w >> 24;
            float oo[3] = {o.x, o.y, o.z};
            float inv[3] = {ix, iy, iz};
            uint32_t hits = 0;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                float bn = tmin, bf = best_t;
                float tn3[3], tf3[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    uint32_t wl = (s >> 1) == 0 ? ql[k].x : ((s >> 1) == 1 ? ql[k].y : ((s >> 1) == 2 ? ql[k].z : ql[k].w));
                    uint32_t wh = (s >> 1) == 0 ? qh[k].x : ((s >> 1) == 1 ? qh[k].y : ((s >> 1) == 2 ? qh[k].z : qh[k].w));
                    uint32_t qlo = (s & 1) ? (wl >> 16) : (wl & 0xffffu);
                    uint32_t qhi = (s & 1) ? (wh >> 16) : (wh & 0xffffu);
                    float plo = org[k] + (float)qlo * step[k];
                    float phi = org[k] + (float)qhi * step[k];
                    float t0 = (plo - oo[k]) * inv[k];
                    float t1 = (phi - oo[k]) * inv[k];
                    bool neg = (oct >> k) & 1u;
                    tn3[k] = neg ? t1 : t0;
                    tf3[k] = neg ? t0 : t1;
                }
                float n_ = fmaxf(fmaxf(tn3[0], tn3[1]), tn3[2]) * kLo;
                float f_ = fminf(fminf(tf3[0], tf3[1]), tf3[2]) * kHi;
                bn = fmaxf(bn, n_);
                bf = fminf(bf, f_);
                if (bn <= bf) hits |= 1u << s;
            }
            uint32_t inner = hits & imask;
            uint32_t leafs = hits & ~imask;
            uint32_t ng = 0;
            while (inner) {
                uint32_t s = __ffs(inner) - 1;
                inner &= inner - 1u;
                ng |= 1u << (s ^ oct);
            }
            uint32_t tm = 0;
            uint32_t meta_lo = w1.z, meta_hi = w1.w;
            while (leafs) {
                uint32_t s = __ffs(leafs) - 1;
                leafs &= leafs - 1u;
                uint32_t m = (s < 4 ? (meta_lo >> (8 * s)) : (meta_hi >> (8 * (s - 4)))) & 0xffu;
                uint32_t off = m & 31u, c = m >> 5;
                tm |= ((1u << c) - 1u) << off;
            }
            g_base = w1.x;
            g_imask = imask;
            g_hits = ng;
            t_base = w1.y;
            t_hits = tm;
        } else if (!t_hits) {
            if (sp == 0) break;
            --sp;
            uint2 e = (sp < ls.cap) ? ls.lds[sp * ls.lds_stride] : ls.spill[(size_t)(sp - ls.cap) * ls.spill_stride];
            g_base = e.x;
            g_imask = e.y & 0xffu;
            g_hits = e.y >> 8;
            continue;
        }
        while (t_hits) {
            uint32_t b = __ffs(t_hits) - 1;
            t_hits &= t_hits - 1u;
            uint32_t ri = t_base + b;
            const uint4 *tp = reinterpret_cast<const uint4 *>(sc.tris + ri);
            uint4 a = tp[0], bq = tp[1], c = tp[2];
            if (COUNT) ++cnt_tris;
            if (SKIP_LIGHTS && c.w != 0u) continue;
            v3 p0(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
            v3 p1(__uint_as_float(bq.x), __uint_as_float(bq.y), __uint_as_float(bq.z));
            v3 p2(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z));
            float t, u, v;
            if (tri_intersect(rp, o, p0, p1, p2, tmin, best_t, &t, &u, &v)) {
                if (ANY) {
                    best_t = t; best_id = a.w; best_rec = ri; bu = u; bv = v;
                    done = true;
                    break;
                }
                if (closer(t, a.w, best_t, best_id)) {
                    best_t = t; best_id = a.w; best_rec = ri; bu = u; bv = v;
                }
            }
        }
        if (ANY && done) break;
    }
    hit.t = best_t;
    hit.u = bu;
    hit.v = bv;
    hit.id = best_id;
    hit.rec = best_rec;
    return best_id != kNoHit;
}

} // namespace fl
