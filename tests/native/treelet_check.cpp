// Host check of the BVH builder's treelet restructuring (fl_treelet.h): a
// binary radix tree (Karras 2012) over the Morton order of synthetic
// primitive boxes, bottom-up sweeps of treelet restructuring in the order the
// device kernel's second-arrival rule gives, then every invariant of the tree:
// each primitive reached exactly once, each internal node once, parent links
// matching child links, node boxes equal to the union of their children,
// primitive counts and subtree costs consistent, and the cost not increased.
// Usage: treelet_check N ROUNDS MODE [boxes.f32]   (MODE 0 uniform cloud,
// 1 sphere shell, 3 all boxes identical, 4 a shell plus a few huge boxes)
// Prints "cost BEFORE AFTER" and "violations V"; exit status 1 on a violation.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <random>
#include <cmath>
#include "fl_treelet.h"
using namespace fl;
static const uint32_t LB = kTreeletLeafBit;
struct Tree {
    int n;
    std::vector<uint32_t> l, r, cnt; std::vector<int> pint, pleaf; std::vector<float> nbox, ncost, tbox;
};
struct HostM {
    Tree *t;
    uint32_t left(uint32_t x) { return t->l[x]; }
    uint32_t right(uint32_t x) { return t->r[x]; }
    void box_of(uint32_t id, float b[6]) {
        const float *s = (id & LB) ? &t->tbox[6 * (id & ~LB)] : &t->nbox[6 * id];
        for (int e = 0; e < 6; ++e) b[e] = s[e];
    }
    float area_of(uint32_t id) { float b[6]; box_of(id, b); return box_half_area(b); }
    float cost(uint32_t id) { return t->ncost[id]; }
    uint32_t count(uint32_t id) { return t->cnt[id]; }
    void set_parent_leaf(uint32_t p, int v) { t->pleaf[p] = v; }
    void set_parent_int(uint32_t p, int v) { t->pint[p] = v; }
    void set_children(uint32_t x, uint32_t a, uint32_t b) { t->l[x] = a; t->r[x] = b; }
    void set_node(uint32_t x, const float b[6], uint32_t c, float cost) {
        for (int e = 0; e < 6; ++e) t->nbox[6 * x + e] = b[e];
        t->cnt[x] = c; t->ncost[x] = cost;
    }
    void set_cost(uint32_t x, float cost) { t->ncost[x] = cost; }
};
static int delta(const std::vector<uint64_t> &k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __builtin_clz((uint32_t)(i ^ j));
    return __builtin_clzll(a ^ b);
}
static uint64_t expand21(uint64_t v) {
    v &= 0x1fffff; v = (v | v << 32) & 0x1f00000000ffffull; v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full; v = (v | v << 4) & 0x10c30c30c30c30c3ull; v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
int main(int argc, char **argv) {
    int n = argc > 1 ? atoi(argv[1]) : 20000; int rounds = argc > 2 ? atoi(argv[2]) : 3; int mode = argc > 3 ? atoi(argv[3]) : 0;
    std::vector<float> filebox;
    if (argc > 4) { FILE *f = fopen(argv[4], "rb"); fseek(f, 0, SEEK_END); long sz = ftell(f); fseek(f, 0, SEEK_SET); filebox.resize(sz / 4); fread(filebox.data(), 4, filebox.size(), f); fclose(f); n = (int)(filebox.size() / 6); mode = 2; }
    std::mt19937 rng(7); std::uniform_real_distribution<float> U(0, 1);
    Tree T; T.n = n; T.tbox.resize(6 * n);
    std::vector<float> cen(3 * n);
    for (int i = 0; i < n; ++i) {
        if (mode == 2) { for (int e = 0; e < 6; ++e) T.tbox[6 * i + e] = filebox[6 * i + e]; for (int e = 0; e < 3; ++e) cen[3 * i + e] = 0.5f * (filebox[6 * i + e] + filebox[6 * i + 3 + e]); continue; }
        float c[3];
        if (mode == 3) { for (int e = 0; e < 3; ++e) { T.tbox[6 * i + e] = 0.f; T.tbox[6 * i + 3 + e] = 1.f; cen[3 * i + e] = 0.5f; } continue; }
        if (mode == 4 && i < 4) { for (int e = 0; e < 3; ++e) { T.tbox[6 * i + e] = -10.f; T.tbox[6 * i + 3 + e] = 10.f; cen[3 * i + e] = 0.3f * (float)(i + 1) - 0.7f; } continue; }
        if (mode == 0) { for (int e = 0; e < 3; ++e) c[e] = U(rng); }
        else { float th = 6.2831853f * U(rng), ph = acosf(2 * U(rng) - 1); c[0] = sinf(ph) * cosf(th); c[1] = sinf(ph) * sinf(th); c[2] = cosf(ph) * (1 + 0.05f * U(rng)); }
        float s = mode == 0 ? 0.01f : 0.004f;
        for (int e = 0; e < 3; ++e) { float a = c[e] - s * U(rng), b = c[e] + s * U(rng); T.tbox[6 * i + e] = a; T.tbox[6 * i + 3 + e] = b; cen[3 * i + e] = 0.5f * (a + b); }
    }
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = 0; i < n; ++i) for (int e = 0; e < 3; ++e) { lo[e] = std::min(lo[e], cen[3 * i + e]); hi[e] = std::max(hi[e], cen[3 * i + e]); }
    std::vector<std::pair<uint64_t, int>> kv(n);
    for (int i = 0; i < n; ++i) {
        uint64_t q[3]; for (int e = 0; e < 3; ++e) q[e] = (uint64_t)std::min(2097151.f, (cen[3 * i + e] - lo[e]) / (hi[e] - lo[e] + 1e-30f) * 2097152.f);
        kv[i] = {(expand21(q[0]) << 2) | (expand21(q[1]) << 1) | expand21(q[2]), i};
    }
    std::sort(kv.begin(), kv.end());
    std::vector<uint64_t> keys(n); std::vector<float> sb(6 * n);
    for (int i = 0; i < n; ++i) { keys[i] = kv[i].first; for (int e = 0; e < 6; ++e) sb[6 * i + e] = T.tbox[6 * kv[i].second + e]; }
    T.tbox = sb;  // primitive boxes in sorted order (leaf index = sorted position)
    int ni = n - 1;
    T.l.resize(ni); T.r.resize(ni); T.pint.assign(ni, -2); T.pleaf.assign(n, -2); T.nbox.resize(6 * ni); T.ncost.assign(ni, 0); T.cnt.resize(ni);
    for (int i = 0; i < ni; ++i) {
        int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) > 0 ? 1 : -1;
        int dmin = delta(keys, n, i, i - d), lmax = 2;
        while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0; for (int t = lmax >> 1; t >= 1; t >>= 1) if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d, dnode = delta(keys, n, i, j), s = 0, t = l;
        do { t = (t + 1) >> 1; if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t; } while (t > 1);
        int gamma = i + s * d + (d < 0 ? d : 0); int lo2 = std::min(i, j), hi2 = std::max(i, j);
        uint32_t L = lo2 == gamma ? (gamma | LB) : gamma, R = hi2 == gamma + 1 ? ((gamma + 1) | LB) : (gamma + 1);
        T.l[i] = L; T.r[i] = R;
        if (L & LB) T.pleaf[L & ~LB] = i; else T.pint[L] = i;
        if (R & LB) T.pleaf[R & ~LB] = i; else T.pint[R] = i;
    }
    T.pint[0] = -1;
    HostM m{&T};
    // refit (post-order by the same bottom-up arrival rule)
    auto sweep = [&](bool restructure) {
        std::vector<int> flags(ni, 0);
        TreeletTable<kTreeletMax> tab;
        std::vector<float> scost(3u << kTreeletMax);
        std::vector<uint8_t> spart(3u << kTreeletMax);
        TreeletTableStrided<3> strided{scost.data() + 1, spart.data() + 1};
        for (int i = 0; i < n; ++i) {
            int node = T.pleaf[i];
            while (node >= 0) {
                if (flags[node]++ == 0) break;
                // as the device kernels do: treelets of at most
                // kTreeletSmall leaves keep their topology; larger ones are
                // optimised in one thread (a strided table, as in LDS) or with
                // the subsets dealt out over 64 lanes, alternately
                if (restructure) {
                    TreeletLeaves<kTreeletMax> t;
                    treelet_form<kTreeletMax>(m, (uint32_t)node, t);
                    if (t.nl <= kTreeletSmall) treelet_keep(m, (uint32_t)node);
                    else if (node & 1) treelet_optimise<kTreeletMax>(m, (uint32_t)node, t, tab, 64u);
                    else treelet_optimise<kTreeletMax>(m, (uint32_t)node, t, strided, 1u);
                }
                else {
                    float b[6] = {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f}, cb[6]; uint32_t c = 0; float cc = 0;
                    for (uint32_t ch : {T.l[node], T.r[node]}) { m.box_of(ch, cb); for (int e = 0; e < 3; ++e) { b[e] = std::min(b[e], cb[e]); b[3 + e] = std::max(b[3 + e], cb[3 + e]); } c += (ch & LB) ? 1 : T.cnt[ch]; cc += (ch & LB) ? 0.f : T.ncost[ch]; }
                    m.set_node(node, b, c, treelet_cost(b, c, cc));
                }
                node = T.pint[node];
            }
        }
    };
    sweep(false);
    double sah0 = T.ncost[0] / box_half_area(&T.nbox[0]);
    for (int r = 0; r < rounds; ++r) sweep(true);
    const double sah1 = T.ncost[0] / box_half_area(&T.nbox[0]);
    printf("cost %.6f %.6f\n", sah0, sah1);
    // invariants
    int bad = 0; std::vector<int> seen_leaf(n, 0), seen_int(ni, 0);
    std::vector<uint32_t> st{0u};
    while (!st.empty()) {
        uint32_t x = st.back(); st.pop_back();
        if (x & LB) { seen_leaf[x & ~LB]++; continue; }
        if (seen_int[x]++) { bad++; continue; }
        float b[6] = {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f}, cb[6]; uint32_t c = 0; double cc = 0;
        for (uint32_t ch : {T.l[x], T.r[x]}) {
            if (ch & LB) { if (T.pleaf[ch & ~LB] != (int)x) bad++; } else { if (T.pint[ch] != (int)x) bad++; }
            m.box_of(ch, cb); for (int e = 0; e < 3; ++e) { b[e] = std::min(b[e], cb[e]); b[3 + e] = std::max(b[3 + e], cb[3 + e]); }
            c += (ch & LB) ? 1 : T.cnt[ch]; cc += (ch & LB) ? 0.f : T.ncost[ch]; st.push_back(ch);
        }
        for (int e = 0; e < 6; ++e) if (b[e] != T.nbox[6 * x + e]) bad++;
        if (c != T.cnt[x]) bad++;
        if (fabs(treelet_cost(b, c, (float)cc) - T.ncost[x]) > 1e-3 * T.ncost[x] + 1e-6) bad++;
    }
    for (int i = 0; i < n; ++i) if (seen_leaf[i] != 1) bad++;
    for (int i = 0; i < ni; ++i) if (seen_int[i] != 1) bad++;
    if (T.cnt[0] != (uint32_t)n) bad++;
    if (sah1 > sah0 * (1.0 + 1e-5)) bad++;
    printf("violations %d\n", bad);
    return bad != 0;
}
