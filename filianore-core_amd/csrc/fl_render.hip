// fl_render.hip -- wavefront spectral path tracer (PathIntegrator::Li for
// every camera sample, integrators/pathintegrator.cpp:12-87, with
// uniform_sample_all_lights/estimate_direct, integrators/integrator.cpp:10-110).
//
// Kernels per pass:  generate -> { extend -> shade -> shadow } x bounces -> resolve
//   generate : camera rays (PhysicalCamera::awaken_ray, cameras/physical.cpp:12-22,
//              pinhole), hero wavelengths, path state init
//   extend   : closest hit of the ray queue (persistent, LDS short stack)
//   shade    : emission (+MIS), NEE light samples -> shadow queue, BSDF
//              sample -> next ray queue, depth limits, Russian roulette
//   shadow   : any hit ignoring emitters (VisibilityEvaluator::unoccluded,
//              illuminants/illuminant.cpp:40-50); occluded NEE terms are zeroed
//   resolve  : hero lanes -> CIE XYZ, per-pixel sums in a fixed order
// Queues are SoA arrays appended with one atomic per wave (ballot + popcount).
// The MIS BSDF ray of estimate_direct is the next extension ray (same
// expectation, one traversal fewer per light per vertex).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fl_wavefront.h"

namespace {

using namespace fl;

constexpr int kBlock = 256;
constexpr int kStack = 16;

__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t *counter) {
    unsigned long long mask = __ballot(want);
    if (mask == 0ull) return 0u;
    uint32_t lane = __lane_id();
    int leader = __ffsll(mask) - 1;
    uint32_t base = 0;
    if (lane == (uint32_t)leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

__device__ __forceinline__ v3 ld3(const float *p) { return v3(p[0], p[1], p[2]); }

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
__global__ void k_zero_f32(float *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// ------------------------------------------------------------- generate ----
__global__ void __launch_bounds__(kBlock) k_generate(RenderArgs a) {
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t nround = (a.n_paths + kBlock - 1) / kBlock * kBlock;
    for (uint32_t base = blockIdx.x * kBlock; base < nround; base += stride) {
        uint32_t sid = base + threadIdx.x;
        bool active = sid < a.n_paths;
        uint32_t x = 0, y = 0, pixel = 0, sample = 0;
        if (active) {
            uint32_t s = sid / a.npix, pl = sid - s * a.npix;
            uint32_t tile = pl >> 6, within = pl & 63u;
            x = (tile % a.tiles_x) * 8u + (within & 7u);
            y = (tile / a.tiles_x) * 8u + (within >> 3);
            active = x < a.cam.W && y < a.cam.H;
            pixel = y * a.cam.W + x;
            sample = a.sample_base + s;
        }
        v3 o, d;
        if (active) {
            uint32_t rb = rng_base(a.seed, pixel, sample, 0xffu);
            float jx = rng_u(rb, 0), jy = rng_u(rb, 1), ul = rng_u(rb, 2);
            float lam0 = a.grid.lo + ul * (a.grid.hi - a.grid.lo);
            if (lam0 >= a.grid.hi) lam0 = a.grid.lo;
            // pixel column px -> s = 1 - (px+jx)/W, row py (top=0) -> t = 1 - (py+jy)/H
            float su = 1.f - ((float)x + jx) / (float)a.cam.W;
            float tv = 1.f - ((float)y + jy) / (float)a.cam.H;
            v3 llc = ld3(a.cam.llc), hor = ld3(a.cam.hor), ver = ld3(a.cam.ver);
            o = ld3(a.cam.o);
            v3 dir = llc + hor * su + ver * tv - o - v3(0.f, 0.f, 0.f);
            d = normalize(dir);
            a.st_beta[sid] = make_float4(1.f, 1.f, 1.f, 1.f);
            a.st_L[sid] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.st_misc[sid] = make_float4(lam0, 1.f, 0.f, __uint_as_float(0u));
            a.st_key[sid] = make_uint2(pixel, sample);
            a.pend[sid] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.pend[a.cap + sid] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else if (sid < a.n_paths) {
            a.st_L[sid] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.st_misc[sid] = make_float4(0.f, 1.f, 0.f, __uint_as_float(PF_DONE));
            a.pend[sid] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.pend[a.cap + sid] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        uint32_t slot = wave_append(active, &a.counts[0]);
        if (active) {
            a.rq_o[0][slot] = make_float4(o.x, o.y, o.z, 0.f);       // tMin = time = 0 (physical.cpp:21)
            a.rq_d[0][slot] = make_float4(d.x, d.y, d.z, kInf);
            a.rq_sid[0][slot] = sid;
        }
    }
}

// --------------------------------------------------------------- extend ----
template <bool COUNT>
__global__ void __launch_bounds__(kBlock) k_extend(RenderArgs a, int cur) {
    __shared__ uint2 stk[kStack * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t nthreads = gridDim.x * kBlock;
    const uint32_t n = a.counts[cur];
    if (gtid == 0) {
        a.counts[cur ^ 1] = 0u;
        a.counts[2] = 0u;
        atomicAdd(&a.stats[0], (unsigned long long)n);
    }
    LaneStack ls;
    ls.lds = stk + threadIdx.x;
    ls.lds_stride = kBlock;
    ls.cap = kStack;
    ls.spill = a.spill + gtid;
    ls.spill_stride = a.spill_stride;
    uint32_t cn = 0, ct = 0, cs = 0;
    for (uint32_t i = gtid; i < n; i += nthreads) {
        float4 ro = a.rq_o[cur][i], rd = a.rq_d[cur][i];
        TraceHit h;
        bool hit = traverse<false, false, COUNT>(a.sc, v3(ro.x, ro.y, ro.z), v3(rd.x, rd.y, rd.z), ro.w, rd.w, ls, h,
                                                 cn, ct, cs);
        a.hits[i] = make_float4(h.t, __uint_as_float(hit ? h.rec : kNoHit), h.u, h.v);
    }
    if (COUNT) {
        atomicAdd(&a.stats[3], (unsigned long long)cn);
        atomicAdd(&a.stats[4], (unsigned long long)ct);
        atomicAdd(&a.stats[5], (unsigned long long)cs);
    }
}

// --------------------------------------------------------------- shadow ----
template <bool COUNT>
__global__ void __launch_bounds__(kBlock) k_shadow(RenderArgs a) {
    __shared__ uint2 stk[kStack * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t nthreads = gridDim.x * kBlock;
    const uint32_t n = a.counts[2];
    if (gtid == 0) atomicAdd(&a.stats[1], (unsigned long long)n);
    LaneStack ls;
    ls.lds = stk + threadIdx.x;
    ls.lds_stride = kBlock;
    ls.cap = kStack;
    ls.spill = a.spill + gtid;
    ls.spill_stride = a.spill_stride;
    uint32_t cn = 0, ct = 0, cs = 0;
    for (uint32_t i = gtid; i < n; i += nthreads) {
        float4 ro = a.sq_o[i], rd = a.sq_d[i];
        TraceHit h;
        bool occluded = traverse<true, true, COUNT>(a.sc, v3(ro.x, ro.y, ro.z), v3(rd.x, rd.y, rd.z), ro.w, rd.w, ls,
                                                    h, cn, ct, cs);
        if (occluded) a.pend[a.sq_tgt[i]] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (COUNT) {
        atomicAdd(&a.stats[6], (unsigned long long)cn);
        atomicAdd(&a.stats[7], (unsigned long long)ct);
        atomicAdd(&a.stats[5], (unsigned long long)cs);
    }
}

// ---------------------------------------------------------------- shade ----
__device__ __forceinline__ void setup_bsdf(const RenderArgs &a, const DevMaterial &m, const Lanes &ln, Bsdf &b) {
    b.kind = m.kind;
    b.ggx.ax = m.ax;
    b.ggx.ay = m.ay;
    b.weight = m.weight;
    b.vndf_pdf = (a.flags & FL_RENDER_VNDF_PDF) != 0;
    const int S = a.grid.stride;
    const float *ta = a.spec + (size_t)m.spec_a * S;
    const float *tb = a.spec + (size_t)m.spec_b * S;
    for (int k = 0; k < 4; ++k) {
        b.r[k] = 0.f; b.t[k] = 0.f; b.eta[k] = 1.f; b.k[k] = 0.f;
        if (m.kind == MAT_LAMBERT) {
            b.r[k] = spec_at(ta, ln, k);
        } else if (m.kind == MAT_CONDUCTOR) {
            b.eta[k] = spec_at(ta, ln, k);
            b.k[k] = spec_at(tb, ln, k);
        } else if (m.kind == MAT_DIELECTRIC) {
            b.r[k] = spec_at(ta, ln, k);
            b.t[k] = spec_at(tb, ln, k);
            float lum = ln.lam[k] * 1e-3f;                      // um
            b.eta[k] = m.cauchy_b != 0.f ? m.eta_bsdf + m.cauchy_b / (lum * lum) : m.eta_bsdf;
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_shade(RenderArgs a, int cur) {
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t n = a.counts[cur];
    const uint32_t nround = (n + kBlock - 1) / kBlock * kBlock;
    const int nxt = cur ^ 1;
    const int S = a.grid.stride;
    for (uint32_t base = blockIdx.x * kBlock; base < nround; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        bool cont = false;
        bool shadow_ok[2] = {false, false};
        float4 sh_o[2], sh_d[2];
        v3 no, nd;
        uint32_t sid = 0;
        if (active) {
            float4 ro = a.rq_o[cur][i], rd = a.rq_d[cur][i];
            sid = a.rq_sid[cur][i];
            float4 hh = a.hits[i];
            float4 beta4 = a.st_beta[sid], L4 = a.st_L[sid], misc = a.st_misc[sid];
            uint2 key = a.st_key[sid];
            float4 pe0 = a.pend[sid], pe1 = a.pend[a.cap + sid];
            float beta[4] = {beta4.x, beta4.y, beta4.z, beta4.w};
            float L[4] = {L4.x + pe0.x + pe1.x, L4.y + pe0.y + pe1.y, L4.z + pe0.z + pe1.z, L4.w + pe0.w + pe1.w};
            float np0[4] = {0.f, 0.f, 0.f, 0.f}, np1[4] = {0.f, 0.f, 0.f, 0.f};
            float lam0 = misc.x, eta_scale = misc.y, prev_pdf = misc.z;
            uint32_t bits = __float_as_uint(misc.w);
            uint32_t bounce = bits & 0xffu;
            uint32_t flags = bits & ~0xffu;
            const uint32_t rec = __float_as_uint(hh.y);
            const float t = hh.x;
            bool alive = rec != kNoHit;
            v3 o(ro.x, ro.y, ro.z), d(rd.x, rd.y, rd.z);
            if (alive) {
                const uint4 *tp = reinterpret_cast<const uint4 *>(a.sc.tris + rec);
                uint4 q0 = tp[0], q1 = tp[1], q2 = tp[2];
                v3 p0(__uint_as_float(q0.x), __uint_as_float(q0.y), __uint_as_float(q0.z));
                v3 p1(__uint_as_float(q1.x), __uint_as_float(q1.y), __uint_as_float(q1.z));
                v3 p2(__uint_as_float(q2.x), __uint_as_float(q2.y), __uint_as_float(q2.z));
                const uint32_t mat = q1.w, light = q2.w;
                v3 p = o + d * t;                                  // ray.point_at_t(t)
                v3 ng = normalize(cross(p1 - p0, p2 - p0));
                Lanes ln = make_lanes(a.grid, lam0);
                const uint32_t rb = rng_base(a.seed, key.x, key.y, bounce);

                // emission: L += beta * le(-dir) at bounce 0 / after specular; MIS otherwise
                if (light) {
                    const DevLight &Lj = a.lights[light - 1];
                    if (dot(ng, -d) > 0.f) {
                        float w = 0.f;
                        if (bounce == 0 || (flags & PF_SPECULAR)) {
                            w = 1.f;
                        } else if (!(flags & PF_PASSED) && prev_pdf > 0.f) {
                            float dl = length(d);
                            float dist2 = (t * t) * (dl * dl);
                            float cosl = fabsf(dot(ng, d)) / dl;
                            float pl = dist2 / (cosl * Lj.area);
                            if (isinf(pl)) pl = 0.f;
                            if (!a.nee_all) pl = pl / (float)a.n_lights;
                            if (pl > 0.f) w = power_heuristic(prev_pdf, pl);
                        }
                        if (w > 0.f) {
                            const float *te = a.spec + (size_t)Lj.em_slot * S;
                            for (int k = 0; k < 4; ++k) L[k] += beta[k] * (spec_at(te, ln, k) * Lj.intensity) * w;
                        }
                    }
                }
                if (flags & PF_LAST) {
                    alive = false;
                } else if (mat == kNoMaterial) {
                    // no BSDF: isect.kindle_ray(ray.dir), bounces-- (pathintegrator.cpp:34-38)
                    v3 perr = vabs(p) * gamma_n(5);
                    no = offset_point_origin(p, perr, ng, d);
                    nd = d;
                    flags |= PF_PASSED;
                    cont = true;
                } else {
                    flags &= ~PF_PASSED;
                    const DevMaterial m = a.mats[mat];
                    Bsdf b;
                    setup_bsdf(a, m, ln, b);
                    // BSDF frame, bsdf.cpp:6-28 (ns = ng: no shading normals)
                    v3 ns = ng, fs, ft;
                    coordinate_system(ns, &fs, &ft);
                    fs = normalize(fs);
                    v3 woW = -d;
                    v3 wo(dot(woW, fs), dot(woW, ft), dot(woW, ns));
                    v3 perr = vabs(p) * gamma_n(5);
                    const bool nonspec = !is_specular_kind(m.kind);
                    // ---- NEE, one sample per light (uniform_sample_all_lights)
                    if (nonspec && a.n_lights > 0) {
                        const int nslots = a.nee_all ? (int)a.n_lights : 1;
                        for (int j = 0; j < nslots; ++j) {
                            uint32_t li = (uint32_t)j;
                            float sel = 1.f;
                            if (!a.nee_all) {
                                li = min((uint32_t)(rng_u(rb, 7) * (float)a.n_lights), a.n_lights - 1u);
                                sel = 1.f / (float)a.n_lights;
                            }
                            const DevLight &Lj = a.lights[li];
                            float ux = rng_u(rb, 3 + 2 * j), uy = rng_u(rb, 4 + 2 * j);
                            // Triangle::sample, triangle.cpp:134-154
                            float uu = sqrtf(ux);
                            float u1 = 1.f - uu, u2 = uy * uu;
                            v3 lp0 = ld3(Lj.p0), lp1 = ld3(Lj.p1), lp2 = ld3(Lj.p2), ln_ = ld3(Lj.n);
                            v3 ps = lp0 * u1 + lp1 * u2 + lp2 * (1.f - u1 - u2);
                            v3 wun = ps - p;
                            float dist2 = length2(wun);
                            if (dist2 == 0.f) continue;
                            v3 wi = normalize(wun);
                            float pdf_l = (1.f / Lj.area) * (length2(p - ps) / absdot(ln_, -wi));
                            if (isinf(pdf_l)) pdf_l = 0.f;
                            pdf_l *= sel;
                            if (!(pdf_l > 0.f)) continue;
                            if (!(dot(ln_, -wi) > 0.f)) continue;           // DiffuseAreaIlluminant::L
                            v3 wil(dot(wi, fs), dot(wi, ft), dot(wi, ns));
                            bool refl = dot(wi, ng) * dot(woW, ng) > 0.f;
                            float f[4];
                            bsdf_eval_nonspec(b, wo, wil, refl, f);
                            float cw = absdot(wi, ns);
                            float pb = bsdf_pdf_nonspec(b, wo, wil);
                            bool black = true;
                            for (int k = 0; k < 4; ++k) { f[k] = f[k] * cw; black = black && f[k] == 0.f; }
                            if (black) continue;
                            float w = power_heuristic(pdf_l, pb);
                            const float *te = a.spec + (size_t)Lj.em_slot * S;
                            float *np = (j == 0) ? np0 : np1;
                            for (int k = 0; k < 4; ++k)
                                np[k] = beta[k] * f[k] * (spec_at(te, ln, k) * Lj.intensity) * w / pdf_l;
                            v3 so = offset_point_origin(p, perr, ng, wun);        // kindle_ray_to
                            sh_o[j] = make_float4(so.x, so.y, so.z, kShadowEps);
                            sh_d[j] = make_float4(wun.x, wun.y, wun.z, 1.f - kShadowEps);
                            shadow_ok[j] = true;
                        }
                    }
                    // ---- BSDF sample (continuation and MIS BSDF ray)
                    BsdfSample bs;
                    float u0 = remap_u0(rng_u(rb, 0)), u1 = rng_u(rb, 1);
                    if (wo.z == 0.f) {
                        bs.pdf = 0.f;
                    } else {
                        lobe_sample(b, wo, u0, u1, bs);
                    }
                    bool ok = bs.pdf != 0.f;
                    v3 wiW;
                    if (ok) {
                        wiW = fs * bs.wi.x + ft * bs.wi.y + ns * bs.wi.z;   // to_world
                        if (!(bs.type & BSDF_SPECULAR)) {
                            bool refl = dot(wiW, ng) * dot(woW, ng) > 0.f;
                            if (!refl) {
                                for (int k = 0; k < 4; ++k) bs.f[k] = 0.f;
                            } else if (b.kind == MAT_CONDUCTOR) {
                                microfacet_f(b, wo, bs.wi, bs.f, &bs.wh);
                            }
                        }
                        bool black = true;
                        for (int k = 0; k < 4; ++k) black = black && bs.f[k] == 0.f;
                        if (black) ok = false;
                    }
                    if (ok) {
                        const uint32_t ty = bs.type;
                        bool brk = (a.depth_d == (int)bounce && (ty & BSDF_DIFFUSE)) ||
                                   (a.depth_r == (int)bounce && (ty & (BSDF_SPECULAR | BSDF_GLOSSY))) ||
                                   (a.depth_t == (int)bounce && (ty & BSDF_TRANSMISSION));
                        if (brk && (ty & BSDF_SPECULAR)) ok = false;
                        if (ok) {
                            if (brk) flags |= PF_LAST;
                            float cw = absdot(wiW, ns);
                            for (int k = 0; k < 4; ++k) beta[k] = beta[k] * bs.f[k] * cw / bs.pdf;
                            if (ty & BSDF_SPECULAR) flags |= PF_SPECULAR; else flags &= ~PF_SPECULAR;
                            if ((ty & BSDF_SPECULAR) && (ty & BSDF_TRANSMISSION)) {
                                float e = m.eta_bsdf;
                                eta_scale *= (dot(woW, ng) > 0.f) ? (e * e) : 1.f / (e * e);
                            }
                            if (bs.secondary_terminated) flags |= PF_SECTERM;
                            no = offset_point_origin(p, perr, ng, wiW);          // kindle_ray
                            nd = wiW;
                            prev_pdf = (ty & BSDF_SPECULAR) ? 0.f : bs.pdf;
                            cont = true;
                            if (!brk) {
                                // Russian roulette, pathintegrator.cpp:77-83
                                float mx = beta[0] * eta_scale;
                                if (!(flags & PF_SECTERM))
                                    for (int k = 1; k < 4; ++k) mx = fmaxf(mx, beta[k] * eta_scale);
                                if (mx < a.rr_threshold && (int)bounce > (int)a.rr_start) {
                                    float q = fmaxf(0.05f, 1.f - mx);
                                    if (rng_u(rb, 2) < q) cont = false;
                                    else for (int k = 0; k < 4; ++k) beta[k] = beta[k] / (1.f - q);
                                }
                            }
                            bounce += 1;
                        }
                    }
                }
            }
            if (!cont) flags |= PF_DONE;
            if (flags & PF_SECTERM) { L[1] = L[2] = L[3] = 0.f; np0[1] = np0[2] = np0[3] = 0.f; np1[1] = np1[2] = np1[3] = 0.f; }
            a.st_L[sid] = make_float4(L[0], L[1], L[2], L[3]);
            a.pend[sid] = make_float4(np0[0], np0[1], np0[2], np0[3]);
            a.pend[a.cap + sid] = make_float4(np1[0], np1[1], np1[2], np1[3]);
            if (!shadow_ok[0]) a.pend[sid] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (!shadow_ok[1]) a.pend[a.cap + sid] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.st_misc[sid] = make_float4(lam0, eta_scale, prev_pdf, __uint_as_float(flags | bounce));
            if (cont) a.st_beta[sid] = make_float4(beta[0], beta[1], beta[2], beta[3]);
        }
        // ---- compaction: one atomic per wave per queue
        uint32_t slot = wave_append(cont, &a.counts[nxt]);
        if (cont) {
            a.rq_o[nxt][slot] = make_float4(no.x, no.y, no.z, kShadowEps);
            a.rq_d[nxt][slot] = make_float4(nd.x, nd.y, nd.z, kInf);
            a.rq_sid[nxt][slot] = sid;
        }
        for (int j = 0; j < 2; ++j) {
            uint32_t ss = wave_append(shadow_ok[j], &a.counts[2]);
            if (shadow_ok[j]) {
                a.sq_o[ss] = sh_o[j];
                a.sq_d[ss] = sh_d[j];
                a.sq_tgt[ss] = (uint32_t)j * a.cap + sid;
            }
        }
    }
}

// -------------------------------------------------------------- resolve ----
__global__ void __launch_bounds__(kBlock) k_resolve(RenderArgs a, const float *cmf) {
    uint32_t pl = blockIdx.x * kBlock + threadIdx.x;
    if (pl >= a.npix) return;
    uint32_t tile = pl >> 6, within = pl & 63u;
    uint32_t x = (tile % a.tiles_x) * 8u + (within & 7u);
    uint32_t y = (tile / a.tiles_x) * 8u + (within >> 3);
    if (x >= a.cam.W || y >= a.cam.H) return;
    const int S = a.grid.stride;
    const float *cx = cmf, *cy = cmf + S, *cz = cmf + 2 * S;
    float sx = 0.f, sy = 0.f, sz = 0.f, qx = 0.f, qy = 0.f, qz = 0.f;
    uint32_t nan_count = 0;
    for (uint32_t s = 0; s < a.spp_pass; ++s) {
        uint32_t sid = s * a.npix + pl;
        if (sid >= a.n_paths) break;
        float4 L4 = a.st_L[sid], p0 = a.pend[sid], p1 = a.pend[a.cap + sid], misc = a.st_misc[sid];
        float L[4] = {L4.x + p0.x + p1.x, L4.y + p0.y + p1.y, L4.z + p0.z + p1.z, L4.w + p0.w + p1.w};
        uint32_t flags = __float_as_uint(misc.w);
        Lanes ln = make_lanes(a.grid, misc.x);
        float X = 0.f, Y = 0.f, Z = 0.f;
        float w = a.xyz_weight;
        for (int k = 0; k < 4; ++k) {
            float wk = (flags & PF_SECTERM) ? (k == 0 ? 4.f * w : 0.f) : w;
            X += wk * spec_at(cx, ln, k) * L[k];
            Y += wk * spec_at(cy, ln, k) * L[k];
            Z += wk * spec_at(cz, ln, k) * L[k];
        }
        if (!isfinite(X) || !isfinite(Y) || !isfinite(Z)) { X = Y = Z = 0.f; ++nan_count; }
        sx += X; sy += Y; sz += Z;
        qx += X * X; qy += Y * Y; qz += Z * Z;
    }
    uint32_t pix = y * a.cam.W + x;
    a.film_sum[3 * pix + 0] += sx;
    a.film_sum[3 * pix + 1] += sy;
    a.film_sum[3 * pix + 2] += sz;
    a.film_sq[3 * pix + 0] += qx;
    a.film_sq[3 * pix + 1] += qy;
    a.film_sq[3 * pix + 2] += qz;
    if (nan_count) atomicAdd(&a.stats[2], (unsigned long long)nan_count);
}

// ----------------------------------------------------------- BSDF queries --
__global__ void k_bsdf_query(const fl_bsdf_query *__restrict__ q, fl_bsdf_result *__restrict__ r, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fl_bsdf_query Q = q[i];
    Bsdf b;
    b.kind = Q.kind;
    b.ggx.ax = Q.alpha_x;
    b.ggx.ay = Q.alpha_y;
    b.weight = 1.f;
    b.vndf_pdf = (Q.flags & FL_RENDER_VNDF_PDF) != 0;
    for (int k = 0; k < 4; ++k) { b.r[k] = Q.r[k]; b.t[k] = Q.t[k]; b.eta[k] = Q.eta[k]; b.k[k] = Q.k[k]; }
    v3 wo(Q.wo[0], Q.wo[1], Q.wo[2]), wi(Q.wi[0], Q.wi[1], Q.wi[2]);
    fl_bsdf_result R;
    float f[4];
    bool refl = wi.z * wo.z > 0.f;  // ng = (0,0,1)
    bsdf_eval_nonspec(b, wo, wi, refl, f);
    for (int k = 0; k < 4; ++k) R.f_eval[k] = f[k];
    R.pdf_eval = bsdf_pdf_nonspec(b, wo, wi);
    BsdfSample s;
    if (wo.z == 0.f) {
        s.pdf = 0.f; s.type = 0; s.wi = v3(0.f, 0.f, 0.f);
        for (int k = 0; k < 4; ++k) s.f[k] = 0.f;
    } else {
        lobe_sample(b, wo, remap_u0(Q.u[0]), Q.u[1], s);
        if (s.pdf == 0.f) {
            s.type = 0;
            for (int k = 0; k < 4; ++k) s.f[k] = 0.f;
        } else if (!(s.type & BSDF_SPECULAR)) {
            bool rf = s.wi.z * wo.z > 0.f;
            if (!rf) for (int k = 0; k < 4; ++k) s.f[k] = 0.f;
            else if (b.kind == MAT_CONDUCTOR) microfacet_f(b, wo, s.wi, s.f, &s.wh);
        }
    }
    for (int k = 0; k < 4; ++k) R.f_sample[k] = s.f[k];
    R.pdf_sample = s.pdf;
    R.wi_sample[0] = s.wi.x; R.wi_sample[1] = s.wi.y; R.wi_sample[2] = s.wi.z;
    R.sampled_type = s.type;
    r[i] = R;
}

template <typename T>
T *grow(void *&p, size_t &cap, size_t bytes, bool &err) {
    if (cap < bytes) {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) { err = true; return nullptr; }
        cap = bytes;
    }
    return (T *)p;
}

} // namespace

// ----------------------------------------------------------- host driver ---
struct RenderBuffers {
    void *p[24] = {};
    size_t cap[24] = {};
};

namespace flx {

static RenderBuffers &bufs_for(fl_ctx_t ctx) {
    static thread_local std::vector<std::pair<fl_ctx_t, RenderBuffers *>> tab;
    for (auto &e : tab)
        if (e.first == ctx) return *e.second;
    tab.push_back({ctx, new RenderBuffers()});
    return *tab.back().second;
}

void release_render_buffers(fl_ctx_t ctx) {
    RenderBuffers &b = bufs_for(ctx);
    for (int i = 0; i < 24; ++i) {
        if (b.p[i]) (void)hipFree(b.p[i]);
        b.p[i] = nullptr;
        b.cap[i] = 0;
    }
}

// PhysicalCamera ctor with flipCameraDirection = true, pinhole (cameras/physical.h:21-43)
static CameraDev make_camera(const fl_camera *c) {
    auto dcross = [](const float a[3], const float b[3], float o[3]) {
        double ax = a[0], ay = a[1], az = a[2], bx = b[0], by = b[1], bz = b[2];
        o[0] = (float)(ay * bz - az * by);
        o[1] = (float)(az * bx - ax * bz);
        o[2] = (float)(ax * by - ay * bx);
    };
    auto norm = [](float v[3]) {
        float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= l; v[1] /= l; v[2] /= l;
    };
    CameraDev cam;
    float pos[3] = {c->position[0], c->position[1], c->position[2]};
    float la[3] = {c->look_at[0], c->look_at[1], c->look_at[2]};
    float up[3] = {c->up[0], c->up[1], c->up[2]};
    float diff[3] = {pos[0] - la[0], pos[1] - la[1], pos[2] - la[2]};
    float fd = std::sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
    float theta = c->vfov_deg * fl::kPi / 180.f;
    float hh = tanf(theta / 2.f);
    float aspect = (float)c->width / (float)c->height;
    float hw = aspect * hh;
    float w[3] = {diff[0], diff[1], diff[2]};
    norm(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] * -1.f;   // flipCameraDirection
    float u[3], v[3];
    dcross(up, w, u);
    norm(u);
    dcross(w, u, v);
    for (int k = 0; k < 3; ++k) w[k] = w[k] * -1.f;   // flip back
    for (int k = 0; k < 3; ++k) {
        cam.o[k] = pos[k];
        cam.llc[k] = pos[k] - (u[k] * hw * fd) - (v[k] * fd * hh) - (w[k] * fd);
        cam.hor[k] = u[k] * 2.f * hw * fd;
        cam.ver[k] = v[k] * 2.f * hh * fd;
    }
    cam.W = c->width;
    cam.H = c->height;
    return cam;
}

int render(fl_scene_s *s, const fl_camera *camera, const fl_render_params *p, fl_film *film, fl_stats *stats) {
    fl_ctx_t ctx = s->ctx;
    hipStream_t st = ctx->stream;
    if (!camera || !p || camera->width == 0 || camera->height == 0 || p->spp == 0)
        return set_error(ctx, FL_E_INVALID, "fl_render: bad camera/params");
    if (p->flags & FL_RENDER_GUIDED) return set_error(ctx, FL_E_UNSUPPORTED, "guided rendering not built");
    (void)hipGetLastError();  // do not report a stale error of an earlier call
    const int mode = p->spectral_mode == FL_SPECTRAL_PARITY ? 1 : 0;
    const SpecTables &tab = s->spec[mode];

    RenderArgs a;
    memset(&a, 0, sizeof(a));
    a.sc = TraceScene{s->d_nodes, s->d_tris};
    a.mats = s->d_mats;
    a.lights = s->d_lights;
    a.n_lights = s->n_lights;
    a.nee_all = s->n_lights <= 2 ? 1u : 0u;
    a.spec = tab.d_tab;
    a.grid = tab.grid;
    a.cam = make_camera(camera);
    const uint32_t W = camera->width, H = camera->height;
    a.tiles_x = (W + 7) / 8;
    const uint32_t tiles_y = (H + 7) / 8;
    a.npix = a.tiles_x * tiles_y * 64u;
    uint32_t max_paths = p->max_paths ? p->max_paths : (1u << 22);
    if (max_paths < a.npix) max_paths = a.npix;
    uint32_t spp_pass = std::max(1u, std::min(p->spp, max_paths / a.npix));
    a.cap = a.npix * spp_pass;
    a.seed = p->seed;
    a.depth_d = p->diffuse_depth;
    a.depth_r = p->reflection_depth;
    a.depth_t = p->transmission_depth;
    a.rr_threshold = p->rr_threshold;
    a.rr_start = p->rr_start;
    a.flags = p->flags;
    a.xyz_weight = (tab.grid.hi - tab.grid.lo) / (4.f * 106.856895f);

    RenderBuffers &B = bufs_for(ctx);
    bool err = false;
    const size_t cap = a.cap;
    int grid_pers = ctx->n_cu * 4;
    size_t spill_per = (size_t)(s->depth > (uint32_t)kStack ? s->depth - kStack + 1 : 1);
    a.spill_stride = (uint32_t)grid_pers * kBlock;
    a.st_beta = grow<float4>(B.p[0], B.cap[0], cap * 16, err);
    a.st_L = grow<float4>(B.p[1], B.cap[1], cap * 16, err);
    a.st_misc = grow<float4>(B.p[2], B.cap[2], cap * 16, err);
    a.st_key = grow<uint2>(B.p[3], B.cap[3], cap * 8, err);
    a.pend = grow<float4>(B.p[4], B.cap[4], 2 * cap * 16, err);
    a.rq_o[0] = grow<float4>(B.p[5], B.cap[5], cap * 16, err);
    a.rq_d[0] = grow<float4>(B.p[6], B.cap[6], cap * 16, err);
    a.rq_sid[0] = grow<uint32_t>(B.p[7], B.cap[7], cap * 4, err);
    a.rq_o[1] = grow<float4>(B.p[8], B.cap[8], cap * 16, err);
    a.rq_d[1] = grow<float4>(B.p[9], B.cap[9], cap * 16, err);
    a.rq_sid[1] = grow<uint32_t>(B.p[10], B.cap[10], cap * 4, err);
    a.hits = grow<float4>(B.p[11], B.cap[11], cap * 16, err);
    a.sq_o = grow<float4>(B.p[12], B.cap[12], 2 * cap * 16, err);
    a.sq_d = grow<float4>(B.p[13], B.cap[13], 2 * cap * 16, err);
    a.sq_tgt = grow<uint32_t>(B.p[14], B.cap[14], 2 * cap * 4, err);
    a.counts = grow<uint32_t>(B.p[15], B.cap[15], 64, err);
    a.stats = grow<unsigned long long>(B.p[16], B.cap[16], 256, err);
    a.film_sum = grow<float>(B.p[17], B.cap[17], (size_t)W * H * 12, err);
    a.film_sq = grow<float>(B.p[18], B.cap[18], (size_t)W * H * 12, err);
    a.spill = grow<uint2>(B.p[19], B.cap[19], (size_t)a.spill_stride * spill_per * 8, err);
    if (err) return set_error(ctx, FL_E_OOM, "render buffers");

    const bool counters = (p->flags & FL_RENDER_COUNTERS) != 0;
    hipEvent_t e0, e1;
    FL_HIP_CHECK(ctx, hipEventCreate(&e0));
    FL_HIP_CHECK(ctx, hipEventCreate(&e1));
    auto blocks = [](size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); };
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks((size_t)W * H * 3)), dim3(kBlock), 0, st, a.film_sum, (size_t)W * H * 3);
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks((size_t)W * H * 3)), dim3(kBlock), 0, st, a.film_sq, (size_t)W * H * 3);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, (uint32_t *)a.stats, 32u);
    FL_HIP_CHECK(ctx, hipEventRecord(e0, st));

    uint32_t done_spp = 0, passes = 0, max_iter_seen = 0, n_launch = 0;
    const bool timing = (p->flags & FL_RENDER_TIMING) != 0;
    hipEvent_t tev[4];
    for (int k = 0; k < 4; ++k) FL_HIP_CHECK(ctx, hipEventCreate(&tev[k]));
    double ms_ext = 0.0, ms_sh = 0.0, ms_shd = 0.0;
    const uint32_t max_iter = (uint32_t)std::max({p->diffuse_depth, p->reflection_depth, p->transmission_depth}) * 2 + 8;
    while (done_spp < p->spp) {
        uint32_t this_spp = std::min(spp_pass, p->spp - done_spp);
        a.spp_pass = this_spp;
        a.n_paths = a.npix * this_spp;
        a.sample_base = p->sample_offset + done_spp;
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, a.counts, 16u);
        hipLaunchKernelGGL(k_generate, dim3(std::min<unsigned>(blocks(a.n_paths), ctx->n_cu * 16)), dim3(kBlock), 0, st, a);
        uint32_t iter = 0;
        for (; iter < max_iter; ++iter) {
            int cur = iter & 1;
            if (timing) FL_HIP_CHECK(ctx, hipEventRecord(tev[0], st));
            if (counters) hipLaunchKernelGGL(k_extend<true>, dim3(grid_pers), dim3(kBlock), 0, st, a, cur);
            else hipLaunchKernelGGL(k_extend<false>, dim3(grid_pers), dim3(kBlock), 0, st, a, cur);
            if (timing) FL_HIP_CHECK(ctx, hipEventRecord(tev[1], st));
            hipLaunchKernelGGL(k_shade, dim3(ctx->n_cu * 8), dim3(kBlock), 0, st, a, cur);
            if (timing) FL_HIP_CHECK(ctx, hipEventRecord(tev[2], st));
            if (counters) hipLaunchKernelGGL(k_shadow<true>, dim3(grid_pers), dim3(kBlock), 0, st, a);
            else hipLaunchKernelGGL(k_shadow<false>, dim3(grid_pers), dim3(kBlock), 0, st, a);
            if (timing) FL_HIP_CHECK(ctx, hipEventRecord(tev[3], st));
            ++n_launch;
            uint32_t next = 0;
            FL_HIP_CHECK(ctx, hipMemcpyAsync(&next, a.counts + (cur ^ 1), 4, hipMemcpyDeviceToHost, st));
            FL_HIP_CHECK(ctx, hipStreamSynchronize(st));
            if (timing) {
                float m0 = 0.f, m1 = 0.f, m2 = 0.f;
                (void)hipEventElapsedTime(&m0, tev[0], tev[1]);
                (void)hipEventElapsedTime(&m1, tev[1], tev[2]);
                (void)hipEventElapsedTime(&m2, tev[2], tev[3]);
                ms_ext += m0; ms_sh += m1; ms_shd += m2;
            }
            if (next == 0) { ++iter; break; }
        }
        max_iter_seen = std::max(max_iter_seen, iter);
        hipLaunchKernelGGL(k_resolve, dim3(blocks(a.npix)), dim3(kBlock), 0, st, a, a.spec);
        done_spp += this_spp;
        ++passes;
    }
    FL_HIP_CHECK(ctx, hipEventRecord(e1, st));
    FL_HIP_CHECK(ctx, hipStreamSynchronize(st));
    FL_HIP_CHECK(ctx, hipGetLastError());
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int k = 0; k < 4; ++k) (void)hipEventDestroy(tev[k]);
    if (film && film->xyz_sum) {
        if (film->width != W || film->height != H) return set_error(ctx, FL_E_INVALID, "film size mismatch");
        FL_HIP_CHECK(ctx, hipMemcpy(film->xyz_sum, a.film_sum, (size_t)W * H * 12, hipMemcpyDeviceToHost));
        if (film->xyz_sqsum)
            FL_HIP_CHECK(ctx, hipMemcpy(film->xyz_sqsum, a.film_sq, (size_t)W * H * 12, hipMemcpyDeviceToHost));
    }
    if (stats) {
        unsigned long long h[10];
        FL_HIP_CHECK(ctx, hipMemcpy(h, a.stats, sizeof(h), hipMemcpyDeviceToHost));
        memset(stats, 0, sizeof(*stats));
        stats->ms_total = ms;
        stats->samples = (uint64_t)W * H * p->spp;
        stats->rays_extension = h[0];
        stats->rays_shadow = h[1];
        stats->nan_samples = h[2];
        stats->nodes_visited = h[3];
        stats->tris_tested = h[4];
        stats->stack_spills = h[5];
        stats->shadow_nodes_visited = h[6];
        stats->shadow_tris_tested = h[7];
        stats->ms_extend = ms_ext;
        stats->ms_shade = ms_sh;
        stats->ms_shadow = ms_shd;
        stats->launches_extend = stats->launches_shade = stats->launches_shadow = n_launch;
        stats->passes = passes;
        stats->max_bounce = max_iter_seen;
        stats->paths_per_pass = a.cap;
        stats->stack_depth = kStack;
    }
    return FL_OK;
}

int bsdf_queries(fl_ctx_t ctx, const fl_bsdf_query *d_q, fl_bsdf_result *d_r, uint32_t n) {
    if (n == 0) return FL_OK;
    hipLaunchKernelGGL(k_bsdf_query, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_q, d_r, n);
    FL_HIP_CHECK(ctx, hipGetLastError());
    return FL_OK;
}

} // namespace flx
