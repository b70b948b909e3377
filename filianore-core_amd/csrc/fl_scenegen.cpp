// fl_scenegen.cpp -- procedural workloads of the render path (host code).
//
// Walls and lights are unit quads of ShapeCreator::create_quad
// (creators/shapecreator.cpp:7-38) placed by an orthonormal frame: the local
// vertex (x, y, 0) maps to c + x*sx*U + y*sy*V; with U x V = W the geometric
// normal (v1-v0)x(v2-v0) of both triangles is -W.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/fl_scenegen.h"

namespace {

// PCG32 (O'Neill 2014, pcg-random.org), the generator the reference vendors
// as src/ext/pcg32.h; used here only to draw scene parameters.
struct Pcg32 {
    uint64_t state, inc;
    explicit Pcg32(uint64_t seed, uint64_t seq = 1) {
        state = 0u;
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    float uniform() { return (float)(next() >> 8) * 0x1p-24f; }
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
};

struct V { float x, y, z; };
V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V operator*(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float len(V a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
V nrm(V a) { float l = len(a); return {a.x / l, a.y / l, a.z / l}; }

struct Builder {
    std::vector<float> v;
    std::vector<uint32_t> idx, mat;
    std::vector<fl_material> mats;
    std::vector<fl_area_light> lights;
    std::vector<float> spd;
    uint32_t add_vertex(V p) {
        v.push_back(p.x); v.push_back(p.y); v.push_back(p.z);
        return (uint32_t)(v.size() / 3 - 1);
    }
    void add_tri(uint32_t a, uint32_t b, uint32_t c, uint32_t m) {
        idx.push_back(a); idx.push_back(b); idx.push_back(c);
        mat.push_back(m);
    }
    uint32_t ntri() const { return (uint32_t)(idx.size() / 3); }
    // create_quad: (1,-1),(-1,-1),(-1,1) and (1,-1),(-1,1),(1,1); returns first tri index
    uint32_t add_quad(V c, V U, V Vv, float sx, float sy, uint32_t m) {
        auto P = [&](float x, float y) { return add_vertex(c + U * (x * sx) + Vv * (y * sy)); };
        uint32_t first = ntri();
        uint32_t a = P(1.f, -1.f), b = P(-1.f, -1.f), d = P(-1.f, 1.f);
        add_tri(a, b, d, m);
        uint32_t a2 = P(1.f, -1.f), b2 = P(-1.f, 1.f), d2 = P(1.f, 1.f);
        add_tri(a2, b2, d2, m);
        return first;
    }
    // quad whose geometric normal is n (unit, axis-aligned or general)
    uint32_t add_quad_n(V c, V n, V U, float sx, float sy, uint32_t m) {
        V W = n * -1.f;
        V Vv = cross(W, U);
        return add_quad(c, U, Vv, sx, sy, m);
    }
    uint32_t add_material(const fl_material &m) {
        mats.push_back(m);
        return (uint32_t)mats.size() - 1;
    }
    void add_light_quad(uint32_t first_tri, uint32_t emission, float intensity) {
        for (uint32_t t = first_tri; t < first_tri + 2; ++t) {
            fl_area_light l;
            memset(&l, 0, sizeof(l));
            l.tri = t;
            l.emission = emission;
            l.intensity = intensity;
            lights.push_back(l);
        }
    }
};

fl_material lambert(float r, float g, float b) {
    fl_material m;
    memset(&m, 0, sizeof(m));
    m.kind = FL_MAT_LAMBERT;
    m.rgb[0] = r; m.rgb[1] = g; m.rgb[2] = b;
    m.weight = 1.f;
    return m;
}
fl_material conductor(float alpha, int metal) {
    fl_material m;
    memset(&m, 0, sizeof(m));
    m.kind = FL_MAT_CONDUCTOR;
    m.roughness = std::sqrt(alpha);   // alpha = roughness^2 (standardsurface.cpp:104-106)
    m.anisotropic = 0.f;
    m.weight = 1.f;
    m.metal = metal;
    return m;
}
fl_material dielectric(float ior, float cauchy_b) {
    fl_material m;
    memset(&m, 0, sizeof(m));
    m.kind = FL_MAT_DIELECTRIC;
    m.rgb[0] = m.rgb[1] = m.rgb[2] = 1.f;
    m.rgb_t[0] = m.rgb_t[1] = m.rgb_t[2] = 1.f;
    m.weight = 1.f;
    m.ior = ior;
    m.cauchy_b = cauchy_b;
    return m;
}
fl_material nomat() {
    fl_material m;
    memset(&m, 0, sizeof(m));
    m.kind = FL_MAT_NONE;
    return m;
}

void blackbody(float T, float *out) {
    const double h = 6.62607015e-34, c = 2.99792458e8, k = 1.380649e-23;
    double mx = 0.0;
    std::vector<double> b(FL_SPD_N);
    for (int i = 0; i < FL_SPD_N; ++i) {
        double l = (360.0 + i) * 1e-9;
        b[i] = (2.0 * h * c * c) / (std::pow(l, 5.0) * (std::exp(h * c / (l * k * (double)T)) - 1.0));
        mx = std::max(mx, b[i]);
    }
    for (int i = 0; i < FL_SPD_N; ++i) out[i] = (float)(b[i] / mx);
}

uint32_t add_blackbody(Builder &B, float T) {
    size_t o = B.spd.size();
    B.spd.resize(o + FL_SPD_N);
    blackbody(T, B.spd.data() + o);
    return (uint32_t)(o / FL_SPD_N);
}

// axis-aligned box (12 tris, outward normals) rotated about y by `ang` (rad)
void add_box(Builder &B, V center, V half, float ang, uint32_t m) {
    float cs = std::cos(ang), sn = std::sin(ang);
    auto rot = [&](V p) { return V{p.x * cs + p.z * sn, p.y, -p.x * sn + p.z * cs}; };
    V ax[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    float h[3] = {half.x, half.y, half.z};
    for (int a = 0; a < 3; ++a) {
        for (int sgn = -1; sgn <= 1; sgn += 2) {
            V n = ax[a] * (float)sgn;
            int b = (a + 1) % 3, c = (a + 2) % 3;
            V U = ax[b];
            V W = n * -1.f;
            V Vv = cross(W, U);
            float sx = h[b];
            float sy = h[c];   // Vv = cross(-n, U) is +-axis c
            V fc = n * h[a];
            // build in local frame then rotate/translate
            auto P = [&](float x, float y) { return B.add_vertex(center + rot(fc + U * (x * sx) + Vv * (y * sy))); };
            uint32_t v0 = P(1.f, -1.f), v1 = P(-1.f, -1.f), v2 = P(-1.f, 1.f);
            B.add_tri(v0, v1, v2, m);
            uint32_t w0 = P(1.f, -1.f), w1 = P(-1.f, 1.f), w2 = P(1.f, 1.f);
            B.add_tri(w0, w1, w2, m);
        }
    }
}

// ------------------------------------------------------------ icosphere ---
struct Mesh {
    std::vector<V> v;
    std::vector<uint32_t> f;
};

const Mesh &icosphere(int level) {
    static std::vector<Mesh> cache(12);
    static std::vector<int> built(12, 0);
    if (built[level]) return cache[level];
    Mesh m;
    const float t = (1.f + std::sqrt(5.f)) / 2.f;
    V base[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                  {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (V &p : base) m.v.push_back(nrm(p));
    // counter-clockwise seen from outside: (v1-v0)x(v2-v0) points outward
    uint32_t F[60] = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2,
                      10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5,
                      2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    m.f.assign(F, F + 60);
    for (int l = 0; l < level; ++l) {
        std::unordered_map<uint64_t, uint32_t> mid;
        mid.reserve(m.f.size() * 2);
        std::vector<uint32_t> nf;
        nf.reserve(m.f.size() * 4);
        auto get = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            V p = nrm((m.v[a] + m.v[b]) * 0.5f);
            m.v.push_back(p);
            uint32_t id = (uint32_t)m.v.size() - 1;
            mid.emplace(key, id);
            return id;
        };
        for (size_t i = 0; i < m.f.size(); i += 3) {
            uint32_t a = m.f[i], b = m.f[i + 1], c = m.f[i + 2];
            uint32_t ab = get(a, b), bc = get(b, c), ca = get(c, a);
            uint32_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), q, q + 12);
        }
        m.f.swap(nf);
    }
    cache[level] = std::move(m);
    built[level] = 1;
    return cache[level];
}

// 3-octave value noise: lattice hash, smoothstep weights, x2 frequency and
// x0.5 amplitude per octave; result in [-1, 1].
float lattice(int x, int y, int z, uint32_t seed) {
    uint32_t h = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ (uint32_t)z * 83492791u ^ seed * 2654435761u;
    h ^= h >> 13; h *= 0x5bd1e995u; h ^= h >> 15;
    return (float)(h & 0xffffffu) / 8388607.5f - 1.f;
}
float value_noise(V p, uint32_t seed) {
    int x0 = (int)std::floor(p.x), y0 = (int)std::floor(p.y), z0 = (int)std::floor(p.z);
    float fx = p.x - x0, fy = p.y - y0, fz = p.z - z0;
    auto ss = [](float t) { return t * t * (3.f - 2.f * t); };
    float wx = ss(fx), wy = ss(fy), wz = ss(fz);
    float r = 0.f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                float w = (dx ? wx : 1.f - wx) * (dy ? wy : 1.f - wy) * (dz ? wz : 1.f - wz);
                r += w * lattice(x0 + dx, y0 + dy, z0 + dz, seed);
            }
    return r;
}
float fbm3(V p, uint32_t seed) {
    float s = 0.f, a = 1.f, norm = 0.f, f = 1.f;
    for (int o = 0; o < 3; ++o) {
        s += a * value_noise(p * f, seed + (uint32_t)o * 101u);
        norm += a;
        a *= 0.5f;
        f *= 2.f;
    }
    return s / norm;
}

void add_sphere(Builder &B, int level, V center, float radius, uint32_t noise_seed, uint32_t m) {
    const Mesh &ic = icosphere(level);
    uint32_t base = (uint32_t)(B.v.size() / 3);
    B.v.reserve(B.v.size() + ic.v.size() * 3);
    for (const V &p : ic.v) {
        float d = 1.f + 0.05f * fbm3(p * 2.5f, noise_seed);   // amplitude 5% of radius
        B.add_vertex(center + p * (radius * d));
    }
    B.idx.reserve(B.idx.size() + ic.f.size());
    for (size_t i = 0; i < ic.f.size(); i += 3) B.add_tri(base + ic.f[i], base + ic.f[i + 1], base + ic.f[i + 2], m);
}

void cornell(Builder &B, const fl_gen_options *o, fl_gen_scene *out) {
    Pcg32 rng_box(o->seed ? o->seed : 1u), rng_col(o->seed ? o->seed + 1 : 2u);
    float rgb[4][3];
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 3; ++k) rgb[i][k] = rng_col.uniform(0.1f, 0.9f);
    uint32_t mw = B.add_material(lambert(rgb[0][0], rgb[0][1], rgb[0][2]));  // floor, ceiling, back
    uint32_t ml = B.add_material(lambert(rgb[1][0], rgb[1][1], rgb[1][2]));  // left
    uint32_t mr = B.add_material(lambert(rgb[2][0], rgb[2][1], rgb[2][2]));  // right
    uint32_t mb = B.add_material(lambert(rgb[3][0], rgb[3][1], rgb[3][2]));  // boxes
    uint32_t mn = B.add_material(nomat());                                   // emitter without BSDF
    B.add_quad_n({0, -1, 0}, {0, 1, 0}, {1, 0, 0}, 1, 1, mw);   // floor, normal +y
    B.add_quad_n({0, 1, 0}, {0, -1, 0}, {1, 0, 0}, 1, 1, mw);   // ceiling, normal -y
    B.add_quad_n({0, 0, -1}, {0, 0, 1}, {1, 0, 0}, 1, 1, mw);   // back, normal +z
    B.add_quad_n({-1, 0, 0}, {1, 0, 0}, {0, 0, 1}, 1, 1, ml);   // left, normal +x
    B.add_quad_n({1, 0, 0}, {-1, 0, 0}, {0, 0, 1}, 1, 1, mr);   // right, normal -x
    const float deg = 3.14159265358979323846f / 180.f;
    float a0 = rng_box.uniform(0.f, 90.f) * deg, s0 = rng_box.uniform(0.8f, 1.2f);
    float a1 = rng_box.uniform(0.f, 90.f) * deg, s1 = rng_box.uniform(0.8f, 1.2f);
    V hA = {0.28f * s0, 0.6f * s0, 0.28f * s0}, hB = {0.28f * s1, 0.3f * s1, 0.28f * s1};
    add_box(B, {-0.35f, -1.f + hA.y, -0.3f}, hA, a0, mb);
    add_box(B, {0.38f, -1.f + hB.y, 0.3f}, hB, a1, mb);
    uint32_t em = add_blackbody(B, 6500.f);
    uint32_t lt = B.add_quad_n({0, 0.99f, 0}, {0, -1, 0}, {1, 0, 0}, 0.3f, 0.3f, mn);
    B.add_light_quad(lt, em, 17.f);
    out->camera = fl_camera{{0.f, 0.f, 3.4f}, {0.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, 40.f, 256, 256};
    out->spp = 16;
    out->depth = 8;
}

void spheres_on_ground(Builder &B, Pcg32 &rng, int n, int level, float span_x, float span_z, int metal) {
    int nx = n > 4 ? 4 : n, nz = (n + 3) / 4;
    uint32_t gi = B.add_material(lambert(0.6f, 0.6f, 0.6f));
    B.add_quad_n({0, 0, 0}, {0, 1, 0}, {1, 0, 0}, 20.f, 20.f, gi);     // 2-tri ground plane
    for (int i = 0; i < n; ++i) {
        int ix = i % 4, iz = i / 4;
        float cx = (nx > 1) ? (-span_x * 0.5f + span_x * ix / (float)(nx - 1)) : 0.f;
        float cz = (nz > 1) ? (-span_z * 0.5f + span_z * iz / (float)(nz - 1)) : 0.f;
        cx += rng.uniform(-0.15f, 0.15f);
        cz += rng.uniform(-0.15f, 0.15f);
        uint32_t m;
        if (rng.uniform() < 0.5f) {
            m = B.add_material(conductor(rng.uniform(0.05f, 0.5f), metal));
        } else {
            float r = rng.uniform(0.1f, 0.9f), g = rng.uniform(0.1f, 0.9f), b = rng.uniform(0.1f, 0.9f);
            m = B.add_material(lambert(r, g, b));
        }
        add_sphere(B, level, {cx, 1.f, cz}, 1.f, (uint32_t)(rng.next() & 0xffffu) + 3u, m);
    }
}

void spheres(Builder &B, const fl_gen_options *o, fl_gen_scene *out) {
    Pcg32 rng(o->seed ? o->seed : 3u);
    int level = o->subdiv >= 0 ? o->subdiv : 7;
    int n = o->n_spheres > 0 ? o->n_spheres : 8;
    spheres_on_ground(B, rng, n, level, 7.5f, 2.5f, 1 /* Copper */);
    uint32_t em = add_blackbody(B, 6500.f);
    uint32_t mn = B.add_material(nomat());
    uint32_t lt = B.add_quad_n({0, 8.f, 0}, {0, -1, 0}, {1, 0, 0}, 3.f, 3.f, mn);
    B.add_light_quad(lt, em, 17.f);
    out->camera = fl_camera{{0.f, 4.f, 9.f}, {0.f, 0.8f, 0.f}, {0.f, 1.f, 0.f}, 40.f, 1024, 1024};
    out->spp = 64;
    out->depth = 16;
}

void guided_room(Builder &B, const fl_gen_options *o, fl_gen_scene *out) {
    Pcg32 rng(o->seed ? o->seed : 3u);
    int level = o->subdiv >= 0 ? o->subdiv : 7;
    int n = o->n_spheres > 0 ? o->n_spheres : 8;
    spheres_on_ground(B, rng, n, level, 7.5f, 2.5f, 1);
    // closed room [-6,6] x [0,6] x [-4,4], inward normals; the ceiling carries a
    // slot of 5% of its area through which an outer quad light shines
    uint32_t mw = B.add_material(lambert(0.7f, 0.7f, 0.7f));
    const float X = 6.f, Yh = 6.f, Z = 4.f;
    B.add_quad_n({0, 0.001f, 0}, {0, 1, 0}, {1, 0, 0}, X, Z, mw);          // floor (just above ground)
    B.add_quad_n({0, Yh * 0.5f, -Z}, {0, 0, 1}, {1, 0, 0}, X, Yh * 0.5f, mw); // back
    B.add_quad_n({0, Yh * 0.5f, Z}, {0, 0, -1}, {1, 0, 0}, X, Yh * 0.5f, mw); // front
    B.add_quad_n({-X, Yh * 0.5f, 0}, {1, 0, 0}, {0, 0, 1}, Z, Yh * 0.5f, mw); // left
    B.add_quad_n({X, Yh * 0.5f, 0}, {-1, 0, 0}, {0, 0, 1}, Z, Yh * 0.5f, mw); // right
    // ceiling (area 12 x 8 = 96) with a slot of 4.8 = 5% of it: x in [-5,-4], z in [-2.4,2.4]
    const float sx0 = -5.f, sx1 = -4.f, sz0 = -2.4f, sz1 = 2.4f;
    auto ceil_rect = [&](float x0, float x1, float z0, float z1) {
        B.add_quad_n({0.5f * (x0 + x1), Yh, 0.5f * (z0 + z1)}, {0, -1, 0}, {1, 0, 0}, 0.5f * (x1 - x0), 0.5f * (z1 - z0), mw);
    };
    ceil_rect(-X, sx0, -Z, Z);
    ceil_rect(sx1, X, -Z, Z);
    ceil_rect(sx0, sx1, -Z, sz0);
    ceil_rect(sx0, sx1, sz1, Z);
    // the outer light: in a closed attic above the ceiling ([-X,X] x [Yh, Yh+Ya]
    // x [-Z,Z], Lambert, inward normals), lying on the attic floor beside the
    // slot and facing up. No point of the room sees it, so the room is lit
    // only by light that bounced in the attic and came down through the slot:
    // light sampling from inside the room finds nothing, and paths have to
    // find the slot themselves (hard indirect transport).
    const float Ya = 3.f;
    uint32_t ma = B.add_material(lambert(0.8f, 0.8f, 0.8f));
    B.add_quad_n({0, Yh + Ya, 0}, {0, -1, 0}, {1, 0, 0}, X, Z, ma);                    // attic top
    B.add_quad_n({0, Yh + Ya * 0.5f, -Z}, {0, 0, 1}, {1, 0, 0}, X, Ya * 0.5f, ma);    // attic back
    B.add_quad_n({0, Yh + Ya * 0.5f, Z}, {0, 0, -1}, {1, 0, 0}, X, Ya * 0.5f, ma);     // attic front
    B.add_quad_n({-X, Yh + Ya * 0.5f, 0}, {1, 0, 0}, {0, 0, 1}, Z, Ya * 0.5f, ma);    // attic left
    B.add_quad_n({X, Yh + Ya * 0.5f, 0}, {-1, 0, 0}, {0, 0, 1}, Z, Ya * 0.5f, ma);    // attic right
    uint32_t em = add_blackbody(B, 6500.f);
    uint32_t mn = B.add_material(nomat());
    uint32_t lt = B.add_quad_n({2.5f, Yh + 0.01f, 0.f}, {0, 1, 0}, {1, 0, 0}, 1.f, 1.5f, mn);
    B.add_light_quad(lt, em, 17.f * 40.f);
    out->camera = fl_camera{{0.f, 3.f, 3.9f}, {0.f, 1.f, 0.f}, {0.f, 1.f, 0.f}, 60.f, 1024, 1024};
    out->spp = 64;
    out->depth = 16;
}

// A second reading of the guided room: the spheres inside a closed
// 12-triangle room [-6,6] x [0,6] x [-4,4] (Lambert, inward normals). The
// ceiling quad stops 0.4 short of the back wall: that strip, 12 x 0.4 = 4.8 = 5% of the ceiling, is the only opening.
// The outer quad light hangs above the strip facing down, so light enters
// only through the strip: it falls on a band of the back wall and floor, and
// the rest of the room is lit indirectly (hard indirect transport). Light
// sampling from most room points hits the ceiling.
void skylight_room(Builder &B, const fl_gen_options *o, fl_gen_scene *out) {
    Pcg32 rng(o->seed ? o->seed : 3u);
    int level = o->subdiv >= 0 ? o->subdiv : 7;
    int n = o->n_spheres > 0 ? o->n_spheres : 8;
    spheres_on_ground(B, rng, n, level, 7.5f, 2.5f, 1);
    uint32_t mw = B.add_material(lambert(0.7f, 0.7f, 0.7f));
    const float X = 6.f, Yh = 6.f, Z = 4.f, slot = 0.4f;
    B.add_quad_n({0, 0.001f, 0}, {0, 1, 0}, {1, 0, 0}, X, Z, mw);          // floor (just above ground)
    B.add_quad_n({0, Yh * 0.5f, -Z}, {0, 0, 1}, {1, 0, 0}, X, Yh * 0.5f, mw); // back
    B.add_quad_n({0, Yh * 0.5f, Z}, {0, 0, -1}, {1, 0, 0}, X, Yh * 0.5f, mw); // front
    B.add_quad_n({-X, Yh * 0.5f, 0}, {1, 0, 0}, {0, 0, 1}, Z, Yh * 0.5f, mw); // left
    B.add_quad_n({X, Yh * 0.5f, 0}, {-1, 0, 0}, {0, 0, 1}, Z, Yh * 0.5f, mw); // right
    // ceiling: z in [-Z + slot, Z]; the strip z in [-Z, -Z + slot] stays open
    B.add_quad_n({0, Yh, 0.5f * slot}, {0, -1, 0}, {1, 0, 0}, X, Z - 0.5f * slot, mw);
    uint32_t em = add_blackbody(B, 6500.f);
    uint32_t mn = B.add_material(nomat());
    uint32_t lt = B.add_quad_n({0.f, Yh + 0.6f, -Z + 0.2f}, {0, -1, 0}, {1, 0, 0}, X, 0.6f, mn);
    B.add_light_quad(lt, em, 17.f * 2.5f);
    out->camera = fl_camera{{0.f, 3.f, 3.9f}, {0.f, 1.f, 0.f}, {0.f, 1.f, 0.f}, 60.f, 1024, 1024};
    out->spp = 64;
    out->depth = 16;
}

void batch_scene(Builder &B, const fl_gen_options *o, fl_gen_scene *out) {
    uint32_t k = o->scene_index;
    Pcg32 rng((o->seed ? o->seed : 5u) * 1000003u + k, 7u);
    // triangle target log-uniform in [1e3, 2e6]
    double tgt = std::exp(std::log(1e3) + rng.uniform() * (std::log(2e6) - std::log(1e3)));
    int best_n = 1, best_l = 0;
    double best_err = 1e30;
    // the closest count of n spheres at subdivision l (+ ground and light
    // quads) that stays inside [1e3, 2e6]
    for (int l = 0; l <= 7; ++l)
        for (int n = 1; n <= 8; ++n) {
            double tris = (double)n * 20.0 * std::pow(4.0, l) + 4.0;
            if (tris < 1e3 || tris > 2e6) continue;
            double err = std::fabs(std::log(tris / tgt));
            if (err < best_err) { best_err = err; best_n = n; best_l = l; }
        }
    spheres_on_ground(B, rng, best_n, best_l, 7.5f, 2.5f, (int)(rng.next() % 9u));
    uint32_t em = add_blackbody(B, 6500.f);
    uint32_t mn = B.add_material(nomat());
    uint32_t lt = B.add_quad_n({0, 8.f, 0}, {0, -1, 0}, {1, 0, 0}, 3.f, 3.f, mn);
    B.add_light_quad(lt, em, 17.f);
    static const uint32_t res[4] = {128, 256, 512, 1024};
    uint32_t r = res[rng.next() % 4u];
    out->camera = fl_camera{{0.f, 4.f, 9.f}, {0.f, 0.8f, 0.f}, {0.f, 1.f, 0.f}, 40.f, r, r};
    out->spp = 16;
    out->depth = 12;
}

void terrain(Builder &B, const fl_gen_options *o, fl_gen_scene *out) {
    Pcg32 rng(o->seed ? o->seed : 6u);
    int gx = o->grid_x > 1 ? o->grid_x : 4096, gz = o->grid_y > 1 ? o->grid_y : 2048;
    const float sizex = 200.f, sizez = 100.f;
    uint32_t mg = B.add_material(lambert(0.45f, 0.4f, 0.3f));
    uint32_t base = (uint32_t)(B.v.size() / 3);
    B.v.reserve(B.v.size() + (size_t)gx * gz * 3);
    for (int j = 0; j < gz; ++j)
        for (int i = 0; i < gx; ++i) {
            float x = -sizex * 0.5f + sizex * i / (float)(gx - 1);
            float z = -sizez * 0.5f + sizez * j / (float)(gz - 1);
            float h = 0.f, a = 6.f, f = 0.02f;
            for (int oct = 0; oct < 6; ++oct) {   // fBm
                h += a * value_noise({x * f, 0.37f * oct, z * f}, 6u + (uint32_t)oct);
                a *= 0.5f;
                f *= 2.f;
            }
            B.add_vertex({x, h, z});
        }
    B.idx.reserve(B.idx.size() + (size_t)(gx - 1) * (gz - 1) * 6);
    for (int j = 0; j + 1 < gz; ++j)
        for (int i = 0; i + 1 < gx; ++i) {
            uint32_t a = base + j * gx + i, b = a + 1, c = a + gx, d = c + 1;
            B.add_tri(a, c, b, mg);   // normal +y
            B.add_tri(b, c, d, mg);
        }
    // 200 dispersive prisms (triangular prisms, 8 tris), Cauchy A=1.5, B=0.0042 um^2
    uint32_t mp = B.add_material(dielectric(1.5f, 0.0042f));
    for (int p = 0; p < 200; ++p) {
        float cx = rng.uniform(-60.f, 60.f), cz = rng.uniform(-30.f, 30.f), s = rng.uniform(0.5f, 2.f);
        float cy = 8.f + rng.uniform(0.f, 6.f);
        float ang = rng.uniform(0.f, 6.2831853f);
        float cs = std::cos(ang), sn = std::sin(ang);
        V tri[3] = {{0.f, 1.f, 0.f}, {-0.866f, -0.5f, 0.f}, {0.866f, -0.5f, 0.f}};
        uint32_t vi[6];
        for (int e = 0; e < 2; ++e)
            for (int q = 0; q < 3; ++q) {
                V lp = {tri[q].x * s, tri[q].y * s, (e ? 1.f : -1.f) * s};
                V wp = {cx + lp.x * cs + lp.z * sn, cy + lp.y, cz - lp.x * sn + lp.z * cs};
                vi[e * 3 + q] = B.add_vertex(wp);
            }
        B.add_tri(vi[0], vi[2], vi[1], mp);   // end caps (outward)
        B.add_tri(vi[3], vi[4], vi[5], mp);
        for (int q = 0; q < 3; ++q) {        // 3 side quads = 6 tris
            int r = (q + 1) % 3;
            B.add_tri(vi[q], vi[r], vi[3 + r], mp);
            B.add_tri(vi[q], vi[3 + r], vi[3 + q], mp);
        }
    }
    // GGX metal boulders: small icospheres, alpha in [0.01, 0.3]
    for (int m = 0; m < 24; ++m) {
        uint32_t mm = B.add_material(conductor(rng.uniform(0.01f, 0.3f), (int)(rng.next() % 9u)));
        add_sphere(B, 3, {rng.uniform(-80.f, 80.f), 10.f + rng.uniform(0.f, 4.f), rng.uniform(-40.f, 40.f)}, rng.uniform(1.f, 3.f),
                   (uint32_t)m + 11u, mm);
    }
    uint32_t em = add_blackbody(B, 5800.f);
    uint32_t mn = B.add_material(nomat());
    uint32_t lt = B.add_quad_n({0, 120.f, 0}, {0, -1, 0}, {1, 0, 0}, 40.f, 40.f, mn);
    B.add_light_quad(lt, em, 40.f);
    uint32_t ci = o->camera_index % 8u;
    float ca = 6.2831853f * (float)ci / 8.f;
    out->camera = fl_camera{{60.f * std::cos(ca), 35.f, 45.f * std::sin(ca)}, {0.f, 5.f, 0.f}, {0.f, 1.f, 0.f}, 45.f, 3840, 2160};
    out->spp = 256;
    out->depth = 24;
}

template <typename T>
T *dup(const std::vector<T> &v) {
    T *p = (T *)malloc(sizeof(T) * (v.size() ? v.size() : 1));
    if (p && !v.empty()) memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

} // namespace

extern "C" {

void fl_gen_default_options(fl_gen_options *opt) {
    memset(opt, 0, sizeof(*opt));
    opt->subdiv = -1;
    opt->n_spheres = -1;
}

void fl_gen_blackbody(float temperature_k, float *spd) { blackbody(temperature_k, spd); }

int fl_gen_config(int config, const fl_gen_options *opt, fl_gen_scene *out) {
    if (!out) return FL_E_INVALID;
    fl_gen_options def;
    fl_gen_default_options(&def);
    const fl_gen_options *o = opt ? opt : &def;
    if (o->subdiv > 9) return FL_E_INVALID;
    memset(out, 0, sizeof(*out));
    Builder B;
    switch (config) {
    case FL_GEN_CORNELL: cornell(B, o, out); break;
    case FL_GEN_SPHERES: spheres(B, o, out); break;
    case FL_GEN_BATCH: batch_scene(B, o, out); break;
    case FL_GEN_GUIDED: guided_room(B, o, out); break;
    case FL_GEN_TERRAIN: terrain(B, o, out); break;
    case FL_GEN_GUIDED_SKYLIGHT: skylight_room(B, o, out); break;
    default: return FL_E_INVALID;
    }
    if (o->width) out->camera.width = o->width;
    if (o->height) out->camera.height = o->height;
    out->vertices = dup(B.v);
    out->n_vertices = (uint32_t)(B.v.size() / 3);
    out->indices = dup(B.idx);
    out->n_triangles = (uint32_t)(B.idx.size() / 3);
    out->tri_material = dup(B.mat);
    out->materials = dup(B.mats);
    out->n_materials = (uint32_t)B.mats.size();
    out->lights = dup(B.lights);
    out->n_lights = (uint32_t)B.lights.size();
    out->emission_spd = dup(B.spd);
    out->n_emission = (uint32_t)(B.spd.size() / FL_SPD_N);
    if (!out->vertices || !out->indices || !out->tri_material || !out->materials) {
        fl_gen_free(out);
        return FL_E_OOM;
    }
    return FL_OK;
}

void fl_gen_free(fl_gen_scene *s) {
    if (!s) return;
    free(s->vertices); free(s->indices); free(s->tri_material);
    free(s->materials); free(s->lights); free(s->emission_spd);
    memset(s, 0, sizeof(*s));
}

void fl_gen_describe(const fl_gen_scene *s, fl_scene_desc *d) {
    memset(d, 0, sizeof(*d));
    d->vertices = s->vertices;
    d->n_vertices = s->n_vertices;
    d->indices = s->indices;
    d->n_triangles = s->n_triangles;
    d->tri_material = s->tri_material;
    d->materials = s->materials;
    d->n_materials = s->n_materials;
    d->lights = s->lights;
    d->n_lights = s->n_lights;
    d->emission_spd = s->emission_spd;
    d->n_emission = s->n_emission;
}

} // extern "C"
