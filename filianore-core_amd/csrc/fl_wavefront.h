// fl_wavefront.h -- data shared by the wavefront render kernels.
#pragma once
#include "fl_internal.h"
#include "fl_trace.h"

namespace fl {

// path flags (high bits of the packed misc word; bounce in the low 8 bits)
enum : uint32_t {
    PF_SPECULAR = 1u << 8,   // last scattering was specular (specularBounce)
    PF_PASSED = 1u << 9,     // last hit was a pass-through (material-less primitive)
    PF_LAST = 1u << 10,      // depth limit reached: count MIS emission, then stop
    PF_SECTERM = 1u << 11,   // secondary wavelengths terminated (dispersion)
    PF_DONE = 1u << 12,
};

// PCG output-permutation hash (Jarzynski & Olano, "Hash Functions for GPU
// Rendering", JCGT 2020) used as a counter-based generator keyed on
// (seed, pixel, sample, bounce, dimension).
FL_HD uint32_t hash32(uint32_t x) {
    uint32_t state = x * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
FL_HD uint32_t rng_base(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t bounce) {
    uint32_t h = hash32((uint32_t)seed ^ pixel);
    h = hash32(h + sample);
    return hash32(h ^ (bounce * 0x9E3779B9u + (uint32_t)(seed >> 32)));
}
FL_HD float rng_u(uint32_t base, uint32_t dim) {
    uint32_t h = hash32(base + dim * 0x632BE5ABu);
    return (float)(h >> 8) * 0x1p-24f;
}

struct CameraDev {
    float o[3], llc[3], hor[3], ver[3];
    uint32_t W, H;
};

struct RenderArgs {
    TraceScene sc;
    const DevMaterial *mats;
    const DevLight *lights;
    uint32_t n_lights;
    uint32_t nee_all;        // 1: one NEE sample per light (uniform_sample_all_lights)
    const float *spec;
    SpecGrid grid;
    CameraDev cam;
    uint32_t npix;           // padded pixels per sample layer (8x8 tiles)
    uint32_t tiles_x;
    uint32_t spp_pass;
    uint32_t sample_base;
    uint32_t n_paths;        // npix * spp_pass
    uint32_t cap;            // path capacity (pend stride)
    uint64_t seed;
    int depth_d, depth_r, depth_t;
    float rr_threshold;
    uint32_t rr_start;
    uint32_t flags;          // FL_RENDER_*
    float xyz_weight;        // (hi-lo)/(4*CIE_Y_integral)
    // path state, indexed by path id
    float4 *st_beta, *st_L, *st_misc;
    uint2 *st_key;
    float4 *pend;            // [2][cap]
    // queues
    float4 *rq_o[2], *rq_d[2];
    uint32_t *rq_sid[2];
    float4 *hits;
    float4 *sq_o, *sq_d;
    uint32_t *sq_tgt;
    uint32_t *counts;        // [0],[1] ray queues, [2] shadow queue
    unsigned long long *stats; // [0] ext rays, [1] shadow rays, [2] nan samples, [3] nodes, [4] tris, [5] spills
    // film
    float *film_sum, *film_sq;
    uint2 *spill;
    uint32_t spill_stride;
    int trav_budget;         // traversal iterations between lane refills
    // learned proposal (guided passes)
    uint32_t guided;         // 1: mix the guide's vMF lobes with BSDF sampling
    float guide_mix;         // probability of sampling the guide
    const __bf16 *gW;        // bf16 weights [5][64][64]
    const float *gP;         // fp32 parameters (biases)
    float *lobes;            // [cap][40]: decoded lobes per extension-queue slot
    float box_lo[3], box_inv[3];   // scene box -> unit cube
    // training records (training passes): 16 floats each
    float *recs;
    uint32_t *rec_count;
    uint32_t rec_cap;
    uint32_t *st_rec;        // per path: index + 1 of its last record, 0 = none
};

} // namespace fl
