// fl_wavefront.h -- data shared by the wavefront render kernels.
#pragma once
#include "fl_guide.h"
#include "fl_internal.h"
#include "fl_trace.h"

namespace fl {

// The guide's decoded lobes of a vertex, as the guided shading kernel keeps
// them in LDS for the rest of that vertex's shading: per lobe the mean
// direction (3), the concentration, the cumulative mixture weight (sampling)
// and the weight times the vMF normalisation (density), one value per lane
// and row (lane-contiguous rows: no bank conflicts).
constexpr int kLobeVals = 6;
constexpr int kLobeRows = kGLobes * kLobeVals;   // 48 rows of 64 lanes

// path flags (high bits of the packed misc word; bounce in the low 8 bits)
enum : uint32_t {
    PF_SPECULAR = 1u << 8,   // last scattering was specular (specularBounce)
    PF_PASSED = 1u << 9,     // last hit was a pass-through (material-less primitive)
    PF_LAST = 1u << 10,      // depth limit reached: count MIS emission, then stop
    PF_SECTERM = 1u << 11,   // secondary wavelengths terminated (dispersion)
    PF_DONE = 1u << 12,
    PF_FRESH = 1u << 13,     // a new camera path: throughput 1, radiance and NEE slots 0,
                             // not yet written to memory (its first shade uses the values)
};

// PCG output-permutation hash (Jarzynski & Olano, "Hash Functions for GPU
// Rendering", JCGT 2020) used as a counter-based generator keyed on
// (seed, pixel, sample, bounce, dimension).
FL_HD uint32_t hash32(uint32_t x) {
    uint32_t state = x * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
// The seed and the pixel are hashed apart before they meet: with seed ^ pixel
// as the first key, seed s + 1 at pixel p ^ s ^ (s + 1) drew exactly what seed
// s drew at pixel p, so renders of nearby seeds (replicates, the two renders
// of a comparison) were the same samples shuffled among neighbouring pixels.
FL_HD uint32_t rng_base(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce) {
    uint32_t h = hash32(hash32((uint32_t)seed) ^ hash32(pixel ^ 0xA511E9B3u));
    h = hash32(h + sample);
    return hash32(h ^ (bounce * 0x9E3779B9u + (uint32_t)(seed >> 32)));
}
FL_HD float rng_u(uint32_t base, uint32_t dim) {
    uint32_t h = hash32(base + dim * 0x632BE5ABu);
    return (float)(h >> 8) * 0x1p-24f;
}

// Three floats in 12 bytes: queued ray directions and shadow-ray origins (a
// shadow ray's t range and an extension ray's t_max follow from the ray's
// kind, so no fourth word is stored).
struct Packed3 {
    float x, y, z;
};

struct CameraDev {
    float o[3], llc[3], hor[3], ver[3];
    uint32_t W, H;
};

// One render item: a (scene, camera, film) triple of a batch. fl_render is a
// batch of one. Path ids, queue slots and film pixels of the items are laid
// out item after item; the per-pass fields change between passes.
struct ItemDev {
    TraceScene sc;
    const DevMaterial *mats;
    const DevLight *lights;
    const float *spec;       // spectral tables of the render's mode
    const StdParams *stds;   // StandardSurface lobe stacks
    const TexScene *tex;     // image textures and uvs (null: none)
    uint32_t n_lights, nee_all;
    uint32_t rgb_basis;      // table slot of the RGB->spectrum basis
    CameraDev cam;
    uint32_t npix, tiles_x;  // padded pixels per sample layer (8x8 tiles), tiles per row
    uint32_t spp_pass, sample_base;   // this pass: samples per pixel, index of the first
    uint32_t path_base, slot_base;    // this pass: first path id, first camera-ray slot
    uint32_t pix_base;                // first pixel of this item's film in the film arrays
    uint32_t ppix_base;               // first padded pixel of this item (resolve's index space)
    uint64_t seed;
};

// Scene data one path needs, from its item.
struct SceneRef {
    TraceScene sc;
    const DevMaterial *mats;
    const DevLight *lights;
    const float *spec;
    const StdParams *stds;
    const TexScene *tex;
    uint32_t n_lights, nee_all;
    uint32_t rgb_basis;
    uint64_t seed;
};

struct RenderArgs {
    const ItemDev *items;    // device table, n_items entries
    uint32_t n_items;
    uint32_t *st_item;       // per path: its item (batches only)
    uint32_t total_pix;      // padded pixels of all items (one sample layer each)
    uint32_t n_cam_rays;     // this pass: camera rays of all items (in-image pixels x samples)
    TraceScene sc;           // item 0 (the whole render when n_items == 1)
    const DevMaterial *mats;
    const DevLight *lights;
    uint32_t n_lights;
    uint32_t nee_all;        // 1: one NEE sample per light (uniform_sample_all_lights)
    uint32_t rgb_basis;      // table slot of the RGB->spectrum basis (normal AOV)
    const StdParams *stds;   // StandardSurface lobe stacks
    const TexScene *tex;     // image textures and uvs (null: none)
    const float *spec;
    SpecGrid grid;
    uint32_t n_paths;        // this pass: path ids of all items
    uint32_t cap;            // path capacity (pend stride)
    uint64_t seed;
    int depth_d, depth_r, depth_t;
    float rr_threshold;
    uint32_t rr_start;
    uint32_t flags;          // FL_RENDER_*
    float xyz_weight;        // (hi-lo)/(4*CIE_Y_integral)
    // path state, indexed by path id
    float4 *st_beta, *st_L, *st_misc;
    float4 *pend;            // [2][cap]
    // queues
    float4 *rq_o[2];         // origin, t_min
    Packed3 *rq_d[2];
    uint32_t *rq_sid[2];
    float2 *hits;            // per extension ray: t, triangle record (kNoHit: miss)
    float2 *hit_uv;          // per extension ray: barycentrics (renders with textures; else null)
    Packed3 *sq_o, *sq_d;
    uint32_t *sq_tgt;
    float4 *sq_occ;          // per shadow ray: its NEE term when occluded (shadowColor); null if all black
    uint32_t *counts;        // [0],[1] ray queues, [2] shadow queue
    unsigned long long *stats; // [0] ext rays, [1] shadow rays, [2] nan samples, [3] nodes, [4] tris, [5] spills
    // films of all items, concatenated (pix_base of each item)
    float *film_sum, *film_sq;
    uint2 *spill;
    uint32_t spill_stride;
    int trav_budget;         // traversal iterations between lane refill checks
    int refill_lanes;        // idle lanes that trigger a refill (or all lanes idle)
    int trav_mode;           // kTravBudget, kTravPrefetch or kTravFlat (fl_render.hip)
    // learned proposal (guided passes)
    uint32_t guided;         // 1: mix the guide's vMF lobes with BSDF sampling
    float guide_mix;         // probability of sampling the guide
    uint32_t guide_bounces;  // vertices of bounce < guide_bounces are guided (all: ~0u)
    const __bf16 *gW;        // bf16 weights [5][64][64]
    const float *gP;         // fp32 parameters (biases)
    float box_lo[3], box_inv[3];   // scene box -> unit cube
    // training records (training passes): 16 floats each
    float *recs;
    uint32_t *rec_count;
    uint32_t rec_cap;
    uint32_t *st_rec;        // per path: index + 1 of its last record, 0 = none
};

// The item of path `sid` (0 in single-item renders, read from the path
// state in batches).
__device__ __forceinline__ uint32_t item_of(const RenderArgs &a, uint32_t sid) {
    return a.n_items > 1 ? a.st_item[sid] : 0u;
}

__device__ __forceinline__ SceneRef scene_of(const RenderArgs &a, uint32_t item) {
    SceneRef r;
    if (a.n_items > 1) {
        const ItemDev &it = a.items[item];
        r.sc = it.sc;
        r.mats = it.mats;
        r.lights = it.lights;
        r.spec = it.spec;
        r.n_lights = it.n_lights;
        r.nee_all = it.nee_all;
        r.rgb_basis = it.rgb_basis;
        r.stds = it.stds;
        r.tex = it.tex;
        r.seed = it.seed;
    } else {
        r.sc = a.sc;
        r.mats = a.mats;
        r.lights = a.lights;
        r.spec = a.spec;
        r.n_lights = a.n_lights;
        r.nee_all = a.nee_all;
        r.rgb_basis = a.rgb_basis;
        r.stds = a.stds;
        r.tex = a.tex;
        r.seed = a.seed;
    }
    return r;
}

// Item owning a position in a concatenated range (path ids or padded
// pixels): the last item whose base is <= v. Bases ascend with the item index.
template <typename F>
__device__ __forceinline__ uint32_t item_search(const RenderArgs &a, uint32_t v, F base) {
    uint32_t lo = 0, hi = a.n_items;   // invariant: base(lo) <= v, answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (base(a.items[mid]) <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

} // namespace fl
