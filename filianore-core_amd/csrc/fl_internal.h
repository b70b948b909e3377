// fl_internal.h -- host-side objects behind the C ABI handles.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/fl_gpu.h"
#include "fl_bvh.h"
#include "fl_shade.h"
#include "fl_standard.h"
#include "fl_texture.h"

struct DevMaterial {        // 48 bytes
    int32_t kind;
    float ax, ay;
    float weight;
    float eta_bsdf;         // BSDF::eta (ksIOR when ks>0 && kt>0, else 1)
    uint32_t spec_a;        // table slot: Lambert R / conductor eta / dielectric R
    uint32_t spec_b;        // table slot: conductor k / dielectric T
    float cauchy_b;         // dielectric Cauchy B (um^2), 0 = scalar eta
    uint32_t tex;           // Lambert: 1 + texture of R, 0 = the table's constant
    uint32_t pad[3];
};

// One illuminant (80 bytes). Area lights (type 0) come first, in
// fl_scene_desc.lights order, then the delta lights.
struct DevLight {
    float p0[3], p1[3], p2[3];   // area: the triangle. point, spot: p0 = posIllum;
                                 // spot: p1 = the cone axis (dirIllum);
                                 // directional: p0 = dirIllum = normalize(-direction)
    float n[3];             // area: normalize(cross(p1-p0, p2-p0)), triangle.cpp:146
    float area;             // area: 0.5*|cross(p1-p0, p2-p0)|, triangle.cpp:130-132
    float intensity;
    uint32_t em_slot;       // emission table slot (the light's color)
    uint32_t tri;           // area: emitting triangle; delta: kNoHit
    uint32_t type;          // 0 = area, else FL_LIGHT_*
    uint32_t shadow_slot;   // shadowColor: 0 = black, else its table slot
    float cos_cone, cos_penumbra;   // spot
};

// Spectral tables of one scene in one grid: [cmf x, y, z, materials (2 each),
// emissions], each `stride` floats.
struct SpecTables {
    fl::SpecGrid grid;
    float *d_tab = nullptr;
    std::vector<float> host;
};

struct RenderState;   // fl_render.hip

struct fl_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    int n_cu = 0;
    // scratch reused across calls (grown on demand)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    void *noalloc_scratch = nullptr;  // the device's fixed traversal scratch (fl_intersect_device_noalloc)
    void *develop_buf = nullptr;     // fl_develop / fl_develop_film staging (grown on demand)
    size_t develop_bytes = 0;
    hipEvent_t ev[16] = {};
    // render working set (queues, path state, film, pinned counters, the
    // captured bounce-loop graph), grown on demand and kept between calls
    RenderState *render_state = nullptr;
};

struct fl_scene_s {
    fl_ctx_t ctx = nullptr;
    uint32_t n_tris = 0, n_verts = 0, n_nodes = 0, depth = 0, n_mats = 0, n_lights = 0;
    uint32_t rgb_basis = 0;         // table slot of the first of the 7 RGB->spectrum basis curves
    bool shadow_colors = false;     // some light has a non-black shadowColor
    bool has_standard = false;      // some material is FL_MAT_STANDARD
    bool has_textures = false;      // some material reads an image texture
    fl::Node8 *d_nodes = nullptr;
    fl::TriRec *d_tris = nullptr;
    uint32_t bvh_width = 8;           // children per wide node (autotuner knob)
    // build inputs kept for rebuilds at another width (12 B per vertex, 20 B per triangle)
    float *d_verts = nullptr;
    uint32_t *d_idx = nullptr, *d_mat = nullptr, *d_light = nullptr;
    DevMaterial *d_mats = nullptr;
    DevLight *d_lights = nullptr;
    fl::StdParams *d_std = nullptr;   // StandardSurface lobe stacks (FL_MAT_STANDARD materials)
    fl::DevTexture *d_tex_info = nullptr;
    uint8_t *d_tex_data = nullptr;
    float2 *d_tri_uv = nullptr;       // 3 per triangle when the scene has uvs
    fl::TexScene *d_texscene = nullptr;   // the three above, for the kernels (null: no textures)
    SpecTables spec[2];     // [FL_SPECTRAL_HERO, FL_SPECTRAL_PARITY]
    double build_ms = 0.0;
    float bmin[3], bmax[3];
    std::vector<DevMaterial> h_mats;
    std::vector<DevLight> h_lights;
    std::vector<fl::StdParams> h_std;
};

// Learned proposal: fp32 master parameters + Adam moments, and the bf16
// copies (row-major W[o][i] for the forward, transposed WT[i][o] for the data
// gradient) that the matrix-core kernels read. Parameter order per matrix k:
// W_k (64 x 64, out-major) then b_k (64).
struct fl_guide_s {
    fl_ctx_t ctx = nullptr;
    fl_guide_config cfg{};
    float *d_params = nullptr;       // kGParams fp32 master
    float *d_m = nullptr, *d_v = nullptr;
    __bf16 *d_w = nullptr;           // kGMats x 64 x 64
    __bf16 *d_wt = nullptr;          // kGMats x 64 x 64, transposed
    float *d_grad = nullptr;         // kGParams (sum over the batch)
    float *d_sq = nullptr;           // per-reduce-block sums of squared gradients
    float *d_partial = nullptr;      // per-block partial gradients
    uint32_t partial_blocks = 0;
    float *d_loss = nullptr;         // per-block partial losses
    float *d_stat = nullptr;         // [0] last loss, [1] gradient norm, [2] diverged flag
    float *d_scratch_in = nullptr;   // staging of host records / inputs
    size_t scratch_in_bytes = 0;
    float *d_scratch_out = nullptr;
    size_t scratch_out_bytes = 0;
    double *d_cdf = nullptr;         // prefix sums of render-record weights
    size_t cdf_bytes = 0;
    void *d_scan_tmp = nullptr;      // hipCUB scan workspace
    size_t scan_tmp_bytes = 0;
    uint32_t *d_draw = nullptr;      // record indices drawn for every step of a training pass
    size_t draw_bytes = 0;
    uint32_t steps = 0;
    float last_loss = 0.f;
    int diverged = 0;
};

namespace flx {

int set_error(fl_ctx_t ctx, int code, const std::string &msg);

#define FL_HIP_CHECK(ctx, call)                                                              \
    do {                                                                                     \
        hipError_t _e = (call);                                                              \
        if (_e != hipSuccess)                                                                \
            return flx::set_error((ctx), FL_E_HIP,                                           \
                                  std::string(#call) + ": " + hipGetErrorString(_e));        \
    } while (0)

// BVH build (fl_build.hip)
int build_bvh(fl_ctx_t ctx, fl_scene_s *s, const float *d_verts, const uint32_t *d_idx,
              const uint32_t *d_mat, const uint32_t *d_light);
// the same build from the scene's kept inputs with `width` children per node
int rebuild_bvh(fl_ctx_t ctx, fl_scene_s *s, uint32_t width);

// spectral preprocessing (fl_spectra.cpp); s->h_std must be filled first
void build_spectral_tables(fl_scene_s *s, const fl_scene_desc *desc, int mode,
                           std::vector<DevMaterial> &mats);
// the lobe stack StandardSurfaceMaterial builds for these parameters
fl::StdParams standard_params(const fl_standard_surface &ss);

// traversal (fl_trace.hip)
// fixed_scratch: the refill cursor and spill area in the library's fixed
// device array (no allocation); stream: nullptr = the context's stream
int trace_rays(fl_ctx_t ctx, fl_scene_s *s, const fl_ray *d_rays, fl_hit *d_hits, uint32_t n, uint32_t flags,
               unsigned long long *d_counters, bool fixed_scratch = false, hipStream_t stream = nullptr);
// the device address of the fixed traversal scratch, looked up once per device
int noalloc_scratch_init(fl_ctx_t ctx);
size_t noalloc_scratch_bytes();

// BSDF queries (fl_render.hip)
int bsdf_queries(fl_ctx_t ctx, const fl_bsdf_query *d_q, fl_bsdf_result *d_r, uint32_t n);
int sampler_draw(fl_ctx_t ctx, uint64_t seed, const uint32_t *d_keys, uint32_t n, uint32_t n_dims, float *d_out);
int bsdf_queries_standard(fl_ctx_t ctx, const fl_standard_surface *m, const fl_standard_query *d_q, fl_bsdf_result *d_r,
                          uint32_t n);

// One (scene, camera, film, spp) of a render; fl_render is a batch of one.
struct HostItem {
    fl_scene_s *scene;
    fl_camera cam;
    fl_film *film;
    uint32_t spp;
};
int render_items(fl_ctx_t ctx, const HostItem *items, uint32_t n, const fl_render_params *p, fl_stats *stats);
// The device film sums of item `item` of the context's most recent render.
int last_film(fl_ctx_t ctx, uint32_t item, const float **d_sum, uint32_t *W, uint32_t *H, uint32_t *spp);
int autotune_items(fl_ctx_t ctx, const HostItem *items, uint32_t n, const fl_render_params *p, fl_tune_result *out);

// Online training of the guide on render records (device, 16 floats each:
// position in the unit box 3, normal 3, wavelength 1, roughness 1, direction
// 3, throughput 1, radiance before 1, path id 1, weight 1, unused 1).
int guide_train_on_records(fl_guide_s *g, const float *d_recs, uint32_t n_recs, uint32_t steps, uint32_t batch,
                           uint64_t seed, hipStream_t st);

} // namespace flx
