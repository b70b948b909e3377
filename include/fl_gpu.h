/*
 * fl_gpu.h -- C ABI of the MI355X render path (libfl_gpu.so).
 *
 * Drop-in boundary for filianore-core's render path.  The reference has no C
 * ABI: its "operator API" is the C++ interface set below, and each entry point
 * here replaces one of them (reference files cited per entry point):
 *
 *   Scene(SceneGeometry, illuminants)        core/scene.h:14, core/scene.cpp:7-19
 *   BVH::initialize_scene_geometry/finalise  bvh/bvh.cpp:7-47           -> fl_scene_create
 *   Scene::intersect / BVH::intersect        core/scene.h:20, bvh/bvh.cpp:49-84
 *   Scene::intersect_p / BVH::intersect_p    core/scene.h:22, bvh/bvh.cpp:86-88 -> fl_intersect*
 *   Integrator::Li (per camera sample)       integrators/integrator.h:17,
 *                                            integrators/pathintegrator.cpp:12-87 -> fl_render
 *   BxDF::evaluate/sample/pdf                shading/bxdfs/bxdf.h:29-33  -> fl_bsdf_query
 *
 * Conventions (kept from the reference where it has one):
 *   - status: 0 = FL_OK, negative = FL_E_*; no C++ exception crosses the ABI;
 *     fl_last_error() gives a message for the last failing call on a context.
 *   - host arrays are caller-owned and copied; handles own all device memory.
 *   - a context owns one HIP stream on one device and is externally
 *     synchronised (one host thread at a time), like the reference's
 *     per-thread Sampler; scenes are immutable after creation and may be
 *     shared by any number of render calls, like Scene (const intersect).
 *   - triangle IDs are the flattened (mesh, triangle) order the reference
 *     gives Embree geometries (bvh/bvh.cpp:12-16,41).
 *   - calls block until results are visible to the host unless an *_device
 *     variant is used (then ordering is on the context stream).
 */
#ifndef FL_GPU_H
#define FL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FL_ABI_VERSION 7

/* ---- status codes ------------------------------------------------------ */
enum {
    FL_OK = 0,
    FL_E_INVALID = -1,     /* bad argument / handle */
    FL_E_HIP = -2,         /* HIP runtime error */
    FL_E_OOM = -3,         /* device allocation failed */
    FL_E_UNSUPPORTED = -4, /* feature not available (e.g. no GPU) */
    FL_E_OVERFLOW = -5,    /* a bounded device structure overflowed */
    FL_E_NUMERIC = -6      /* divergence / NaN detected where it is fatal */
};

typedef struct fl_ctx_s *fl_ctx_t;
typedef struct fl_scene_s *fl_scene_t;
typedef struct fl_guide_s *fl_guide_t;

/* ---- materials --------------------------------------------------------- */
/* Material kinds: the StandardSurface branches and LambertMaterial that the
 * render path implements.  FL_MAT_NONE is a primitive without material: the
 * reference lets such hits pass through (pathintegrator.cpp:33-38). */
enum {
    FL_MAT_NONE = 0,
    FL_MAT_LAMBERT = 1,    /* LambertMaterial, shading/materials/lambert.cpp:9-17 */
    FL_MAT_CONDUCTOR = 2,  /* StandardSurface WavelengthDependentMetallic branch,
                              standardsurface.cpp:101-125 (kdweight = 0) */
    FL_MAT_DIELECTRIC = 3, /* StandardSurface ks>0, kt>0 branch: FresnelSpecularBXDF,
                              standardsurface.cpp:77-81,113-117 */
    FL_MAT_STANDARD = 4    /* the whole StandardSurfaceMaterial lobe stack
                              (standardsurface.cpp:72-159), parameters in
                              fl_scene_desc.standard_surfaces[fl_material.standard] */
};

/* StandardSurfaceMaterial's parameters (shading/materials/standardsurface.h:
 * constant textures; colours are reflectance RGB, turned into spectra by
 * from_reflectance_rgb and clamped, as an RGB texture would be). */
enum { FL_FRESNEL_DIELECTRIC = 0, FL_FRESNEL_METALLIC = 1, FL_FRESNEL_WAVELENGTH_METALLIC = 2 };

typedef struct {
    float kd_weight, kd[3], kd_roughness;                  /* diffuse */
    float ks_weight, ks[3], ks_roughness, ks_anisotropic;  /* specular */
    int32_t fresnel_type;                                  /* FL_FRESNEL_* */
    float ks_ior;
    int32_t metal;                                         /* forge metal for the wavelength-dependent metallic */
    float sheen_weight, sheen_color[3], sheen_roughness;   /* sheen */
    float kt_weight, kt[3];                                /* transmission */
    float coat_weight, coat[3], coat_ior, coat_gloss;      /* clearcoat */
    float thin_film_thickness, thin_film_ior;              /* thin film (nm) */
    int32_t thin_walled;
    /* image textures (ImageMapTexture) in place of the constant colours kd,
     * ks, sheen_color, kt, coat: 0 = the constant, else 1 + an index into
     * fl_scene_desc.textures */
    uint32_t kd_texture, ks_texture, sheen_texture, kt_texture, coat_texture;
} fl_standard_surface;

typedef struct {
    int32_t kind;        /* FL_MAT_* */
    float rgb[3];        /* Lambert kd, or dielectric ks (reflectance RGB) */
    float rgb_t[3];      /* dielectric kt (reflectance RGB) */
    float roughness;     /* conductor ksroughness: alpha = roughness^2 (aspect 1) */
    float anisotropic;   /* conductor ksanisotropic */
    float weight;        /* Lambert: unused (1); conductor: ksweight */
    int32_t metal;       /* forge metal 0..8 (metallurgy.cpp:42-75) */
    float ior;           /* dielectric ksIOR; with cauchy_b != 0: Cauchy A */
    float cauchy_b;      /* Cauchy B [um^2]; 0 = scalar eta as in the reference */
    uint32_t standard;   /* FL_MAT_STANDARD: index into fl_scene_desc.standard_surfaces */
    uint32_t texture;    /* Lambert: an ImageMapTexture for kd, 1 + index into
                            fl_scene_desc.textures (0 = the constant rgb) */
    /* A reflectance given as a spectrum instead of reflectance RGB, as a
     * ConstantTexture<PrincipalSpectrum> built by from_spd would hold it
     * (color/spectrumoperations.cpp:88-103; shading/textures/constant.h:9-20):
     * 0 = the rgb above, else 1 + a row of fl_scene_desc.emission_spd,
     * clamped at 0 like the materials clamp their spectra. spd_r stands for
     * the Lambert kd or the dielectric ks, spd_t for the dielectric kt. */
    uint16_t spd_r, spd_t;
} fl_material;

/* One DiffuseAreaIlluminant over one triangle (illuminants/diffusearea.cpp). */
typedef struct {
    uint32_t tri;        /* emitting triangle ID */
    uint32_t emission;   /* index into fl_scene_desc.emission_spd */
    float intensity;     /* DiffuseAreaIlluminant::intensity */
    uint32_t shadow_emission; /* Illuminant::shadowColor, the light's value at an occluded
                                 sample (integrator.cpp:47-50): 0 = black, else
                                 1 + index into emission_spd */
} fl_area_light;

/* Delta illuminants (illuminants/point.cpp, spot.cpp, directional.cpp). They
 * are sampled by NEE only, with pdf 1 and no MIS (is_delta_illuminant,
 * illuminant.cpp:6-8; integrator.cpp:53-54,62). Point and spot lights fall
 * off with the squared distance: Illuminant::evaluate_decay_rate falls
 * through every case to d^2 whatever the decay rate (illuminant.cpp:20-38). */
enum {
    FL_LIGHT_POINT = 1,        /* PointIlluminant */
    FL_LIGHT_SPOT = 2,         /* SpotIlluminant */
    FL_LIGHT_DIRECTIONAL = 3   /* DirectionalIlluminant */
};

typedef struct {
    int32_t type;              /* FL_LIGHT_* */
    float position[3];         /* point, spot: lightToWorld applied to the origin (posIllum) */
    float direction[3];        /* spot: the cone axis, lightToWorld.vector_transform(dirIllum);
                                  directional: lightToWorld.vector_transform(dirIllum), the way
                                  the light travels (the light is at -direction) */
    uint32_t emission;         /* color: index into fl_scene_desc.emission_spd */
    float intensity;
    float cos_cone;            /* spot: cos of the (half) cone angle; no light outside it */
    float cos_penumbra;        /* spot: cos of the (half) penumbra angle; full light inside it */
    uint32_t shadow_emission;  /* shadowColor: 0 = black, else 1 + index into emission_spd */
} fl_delta_light;

#define FL_SPD_N 471     /* emission SPDs are sampled on 360..830 nm at 1 nm */

/* An ImageMapTexture (shading/textures/imagemap.cpp): 8-bit RGB, 3 bytes per
 * texel, rows top to bottom (as stbi_load returns them). evaluate looks up
 * the texel at i = int(u W), j = int((1 - v) H - 0.001), both clamped to the
 * image (the reference clamps j into i, imagemap.cpp:54-55; here j is
 * clamped), and returns from_reflectance_rgb(rgb / 255), clamped at 0 as the
 * materials clamp their spectra. A texture with no data (rgb NULL or a zero
 * size) is black, as an image that failed to load is. */
typedef struct {
    uint32_t width, height;
    const uint8_t *rgb;          /* width * height * 3 bytes */
} fl_texture;

typedef struct {
    const float *vertices;       /* n_vertices * 3 floats */
    uint32_t n_vertices;
    const uint32_t *indices;     /* n_triangles * 3 */
    uint32_t n_triangles;
    const uint32_t *tri_material;/* n_triangles material indices (NULL: all 0) */
    const fl_material *materials;
    uint32_t n_materials;
    const fl_area_light *lights;
    uint32_t n_lights;
    const float *emission_spd;   /* n_emission * FL_SPD_N */
    uint32_t n_emission;
    const fl_delta_light *delta_lights; /* sampled after the area lights, in this order */
    uint32_t n_delta_lights;
    const fl_standard_surface *standard_surfaces;
    uint32_t n_standard_surfaces;
    const fl_texture *textures;
    uint32_t n_textures;
    /* per-vertex texture coordinates (n_vertices * 2 floats), interpolated
     * with the hit's barycentric weights as Triangle::intersect does
     * (shapes/triangle.cpp:66-73); NULL: every hit has uv = (0, 0), as on the
     * reference's Embree path (bvh/bvh.cpp:69-71) */
    const float *uvs;
    /* children per BVH node, 2..8 (0: 8). Nodes always use the 8-slot,
     * 128-byte format; a smaller width leaves slots empty and makes the tree
     * deeper. fl_autotune may rebuild the scene at another width. */
    uint32_t bvh_width;
} fl_scene_desc;

typedef struct {
    uint32_t n_triangles, n_nodes, n_leaf_tris, depth;
    uint64_t node_bytes, tri_bytes;
    double build_ms;             /* device time of the LBVH build */
    float bounds_min[3], bounds_max[3];
} fl_scene_info;

/* ---- rays and hits (Scene::intersect) ----------------------------------- */
typedef struct { float ox, oy, oz, tmin, dx, dy, dz, tmax; } fl_ray; /* 32 B */
typedef struct { float t; uint32_t tri; float u, v; } fl_hit;        /* 16 B */
#define FL_NO_HIT 0xffffffffu

enum {
    FL_CLOSEST_HIT = 0,
    FL_ANY_HIT = 1,          /* shadow query: any hit in (tmin, tmax) */
    FL_SKIP_LIGHTS = 2,      /* any-hit ignores emitting triangles, like
                                VisibilityEvaluator::unoccluded (illuminant.cpp:40-50) */
    FL_COUNTERS = 4          /* count nodes/triangles visited (slower) */
};

/* ---- camera / render ---------------------------------------------------- */
typedef struct {
    float position[3], look_at[3], up[3];
    float vfov_deg;
    uint32_t width, height;
} fl_camera;

enum {
    FL_SPECTRAL_HERO = 0,    /* 4 hero wavelengths, 360..830 nm, continuous spectra */
    FL_SPECTRAL_PARITY = 1   /* 4 hero wavelengths, 400..700 nm, the reference's 60 bins */
};

enum {
    FL_RENDER_GUIDED = 1,        /* use the learned vMF proposal (needs params.guide) */
    FL_RENDER_COUNTERS = 2,      /* per-kernel counters */
    FL_RENDER_VNDF_PDF = 4,      /* GGX: matched VNDF pdf instead of the reference's D*cos */
    FL_RENDER_GRAPH = 8,         /* capture the bounce loop in a hipGraph */
    FL_RENDER_TRAIN_GUIDE = 16,  /* train the guide online on this render's samples */
    FL_RENDER_SORT_SHADING = 32, /* order each shading block's vertices by material kind first
                                    (an autotuner knob; off by default since ABI 7: slower on
                                    every scene measured) */
    FL_RENDER_TIMING = 64,       /* per-kernel HIP-event timing (ms_* fields of fl_stats) */
    FL_RENDER_NORMALS = 128,     /* NormalIntegrator (integrators/normalintegrator.cpp:11-26):
                                    the first hit's normal through from_reflectance_rgb, black
                                    on a miss; no light transport */
    FL_RENDER_NEE_ONE = 256      /* direct lighting from one uniformly chosen light per vertex
                                    (its estimate divided by the selection probability) also in
                                    scenes of 1 or 2 lights, which otherwise take one sample
                                    per light as uniform_sample_all_lights does
                                    (integrators/integrator.cpp:10-23); scenes of more lights
                                    always sample one. Same expectation, half the shadow rays
                                    of a two-triangle emitter */
};

typedef struct {
    uint32_t spp;
    int32_t diffuse_depth, reflection_depth, transmission_depth; /* RenderParams */
    float rr_threshold;          /* PathIntegrator::rrThreshold (default 1) */
    uint32_t rr_start;           /* RR when bounces > rr_start (reference: 3) */
    uint32_t spectral_mode;      /* FL_SPECTRAL_* */
    uint32_t flags;              /* FL_RENDER_* */
    uint64_t seed;
    uint32_t sample_offset;      /* progressive rendering: index of first sample */
    uint32_t max_paths;          /* paths in flight per pass (0: auto) */
    uint32_t stack_depth;        /* LDS short-stack entries per lane (0: auto = 8) */
    uint32_t trav_budget;        /* traversal iterations between lane refills (0: auto) */
    uint32_t shade_waves;        /* shading kernel variant: min waves/SIMD (2, 3 or 4; 0: 3, or 2 with
                                    standard surfaces or textures) */
    uint32_t trav_waves;         /* traversal variant: min waves/SIMD (0: auto); valid
                                    (stack_depth, trav_waves): (8,0) (12,0) (8,6) (8,8) (6,8) (16,0) */
    fl_guide_t guide;            /* guiding network when FL_RENDER_GUIDED */
    float guide_mix;             /* proposal probability of the guide (0.5) */
    float guide_train_fraction;  /* with TRAIN_GUIDE: fraction of spp used to train (0: 0.15) */
    uint32_t refill_lanes;       /* traversal: idle lanes of a wave that trigger a refill (0: auto,
                                    24 with rays loaded ahead, 16 with budgeted refill) */
    uint32_t trav_prefetch;      /* traversal mode: 1 = persistent, each lane's next ray loaded
                                    ahead, the wave yields once refill_lanes lanes are done;
                                    2 = persistent, refill after trav_budget iterations; 3 = one ray
                                    per thread; 0 = auto, per pass: scenes below 65,536 triangles
                                    3 up to 16 M paths per pass, else 1; larger scenes 1 from 4 M
                                    paths per pass on, else 2 */
    uint32_t guide_bounces;      /* guided renders: the guide serves vertices of bounce < guide_bounces
                                    (deeper vertices sample the BSDF alone); 0 = every vertex */
} fl_render_params;

/* Film: per-pixel sums of linear CIE XYZ (3 floats per pixel) and of its
 * square, over `spp` samples.  The pixel estimate is sum/spp. */
typedef struct {
    float *xyz_sum;              /* width*height*3 host floats (NULL: keep on device) */
    float *xyz_sqsum;            /* width*height*3 or NULL */
    uint32_t width, height;
} fl_film;

typedef struct {
    double ms_total, ms_generate, ms_extend, ms_shade, ms_shadow, ms_resolve, ms_guide;
    uint64_t samples, rays_extension, rays_shadow;
    uint64_t nodes_visited, tris_tested;   /* with FL_RENDER_COUNTERS: all rays traced */
    uint64_t shadow_nodes_visited, shadow_tris_tested;   /* ... of which any-hit shadow rays */
    uint32_t launches_extend, launches_shade, launches_shadow, pad0;
    uint64_t nan_samples, stack_spills;
    uint32_t passes, max_bounce;
    uint32_t paths_per_pass, stack_depth;
    double ms_infer;                /* 0: guided renders evaluate the network inside the shading
                                       kernel, whose time is ms_shade */
    uint64_t guide_queries;         /* network evaluations (guided hits) */
} fl_stats;

/* ---- context / scene ---------------------------------------------------- */
int fl_ctx_create(int device, fl_ctx_t *out);
void fl_ctx_destroy(fl_ctx_t ctx);
const char *fl_last_error(fl_ctx_t ctx);
int fl_abi_version(void);
int fl_device_count(void);
/* HBM copy probe (no counterpart in the reference): a plain streaming kernel
 * (k_copy_probe) copies `bytes` (rounded down to 16 B) from one device buffer
 * to another `iters` times; *gbps = (read + write bytes) / time. It is the
 * known-size reference point for memory-counter calibration and a measured
 * bandwidth to set beside the published peak. */
int fl_probe_copy(fl_ctx_t ctx, uint64_t bytes, uint32_t iters, double *gbps);

/* Copies the geometry, builds the 8-wide quantized LBVH on the device. */
int fl_scene_create(fl_ctx_t ctx, const fl_scene_desc *desc, fl_scene_t *out);
void fl_scene_destroy(fl_scene_t scene);
int fl_scene_get_info(fl_scene_t scene, fl_scene_info *out);
/* Copies BVH nodes (128 B each) and triangle records (48 B each) to host
 * buffers sized from fl_scene_get_info (for structural checks). */
int fl_scene_download_bvh(fl_scene_t scene, void *nodes, size_t node_bytes,
                          void *tris, size_t tri_bytes);

/* Closest/any hit for n host rays (replaces Scene::intersect/intersect_p). */
int fl_intersect(fl_scene_t scene, const fl_ray *rays, fl_hit *hits, uint32_t n,
                 uint32_t flags, uint64_t *counters /* [nodes, tris] or NULL */);
/* Same on device-resident arrays, ordered on the context stream. */
int fl_intersect_device(fl_scene_t scene, const fl_ray *d_rays, fl_hit *d_hits,
                        uint32_t n, uint32_t flags);
/* fl_intersect_device without any allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). The
 * traversal's refill cursor and stack spill area live in a fixed 16 MiB
 * device array of the library (one per GPU, held until the process ends)
 * instead of a buffer the context grows; a deep BVH runs fewer persistent
 * blocks so that its spill area fits (FL_E_OVERFLOW if not even one block
 * fits). Results are those of fl_intersect_device bit for bit. stream: a
 * hipStream_t to order the work on (NULL: the context's stream). Every call
 * on one GPU shares that array: calls that overlap on one GPU race, so order
 * them on one stream. FL_COUNTERS is rejected. */
int fl_intersect_device_noalloc(fl_scene_t scene, const fl_ray *d_rays, fl_hit *d_hits,
                                uint32_t n, uint32_t flags, void *stream);

/* Renders `params->spp` samples per pixel (PathIntegrator::Li for every
 * camera sample).  film may be NULL (result stays on the device). */
int fl_render(fl_scene_t scene, const fl_camera *camera, const fl_render_params *params,
              fl_film *film, fl_stats *stats);

/* ---- batched rendering -------------------------------------------------- */
typedef struct {
    fl_scene_t scene;
    fl_camera camera;
    fl_film *film;
    uint32_t spp;
    uint32_t pad;
} fl_batch_item;

/* Renders many (scene, camera) pairs of uneven size in one persistent launch
 * per kernel stage, with the bounce loop captured in a hipGraph. */
int fl_render_batch(fl_ctx_t ctx, const fl_batch_item *items, uint32_t n,
                    const fl_render_params *params, fl_stats *stats);

/* ---- autotuner -------------------------------------------------------------- */
/* Probe renders of the same job (the given spp, at most 64) choose, one knob
 * after the other, the values that minimise device time: BVH node width (the
 * scenes are rebuilt, and left at the chosen width), paths in flight per pass
 * (queue chunk), traversal variant (LDS stack depth x minimum waves per
 * SIMD), traversal steps between lane refills, refill threshold, shading sort
 * on/off, shading-kernel occupancy. No counterpart in the reference. */
typedef struct {
    uint32_t max_paths, stack_depth, trav_waves, trav_budget, shade_waves;
    uint32_t sort_shading;        /* 1: per-block material sort on (FL_RENDER_SORT_SHADING) */
    double ms_best, ms_default;   /* device ms of one render with the chosen and the default values */
    uint32_t n_tried;             /* probe configurations measured */
    uint32_t refill_lanes;        /* traversal refill threshold (0: the mode's default) */
    uint32_t node_width;          /* BVH children per node the scenes now have */
    uint32_t trav_prefetch;       /* traversal refill mode (fl_render_params.trav_prefetch) */
} fl_tune_result;

int fl_autotune(fl_scene_t scene, const fl_camera *camera, const fl_render_params *params,
                fl_tune_result *out);
int fl_autotune_batch(fl_ctx_t ctx, const fl_batch_item *items, uint32_t n, const fl_render_params *params,
                      fl_tune_result *out);

/* ---- learned proposal (vMF mixture MLP) ----------------------------------- */
typedef struct {
    uint32_t hidden;        /* 64 */
    uint32_t layers;        /* 4 hidden layers */
    uint32_t lobes;         /* 8 vMF lobes */
    uint32_t bands;         /* 8 frequency bands per position axis */
    float lr;               /* Adam learning rate */
    float beta1, beta2, eps;
    uint32_t batch;         /* training batch (samples per step) */
    uint64_t seed;
    float scene_min[3], scene_max[3]; /* position normalisation box (fl_render uses the scene's bounds) */
    uint32_t steps_per_pass;  /* Adam steps after each training pass of a guided render (128) */
    uint32_t train_rounds;    /* training passes of doubling size over the training samples (0: 3) */
} fl_guide_config;

int fl_guide_create(fl_ctx_t ctx, const fl_guide_config *cfg, fl_guide_t *out);
void fl_guide_destroy(fl_guide_t guide);
/* One Adam step on n training records (device or host), returns the loss. */
int fl_guide_train_step(fl_guide_t guide, const float *records, uint32_t n, int records_on_device,
                        float *loss);
/* Forward pass: inputs n x 64 floats (encoded), outputs n x 40 floats (raw head). */
int fl_guide_forward(fl_guide_t guide, const float *inputs, float *outputs, uint32_t n);
/* Parameter access (fp32 master weights, then Adam m, v), for checks/checkpoints. */
int fl_guide_get_params(fl_guide_t guide, float *params, uint32_t n_params, uint32_t *n_out);
int fl_guide_set_params(fl_guide_t guide, const float *params, uint32_t n_params);
/* Gradient of the loss on n records w.r.t. all parameters (no update). */
int fl_guide_gradient(fl_guide_t guide, const float *records, uint32_t n, float *grad,
                      uint32_t n_params, float *loss);
int fl_guide_stats(fl_guide_t guide, uint32_t *steps, float *last_loss, int *diverged);

/* ---- BSDF queries (BxDF::evaluate/sample/pdf on the device) ---------------- */
typedef struct {
    int32_t kind;                 /* FL_MAT_* */
    float alpha_x, alpha_y;       /* GGX */
    float eta[4], k[4];           /* conductor eta/k per lane; dielectric: eta[0] = ksIOR */
    float r[4], t[4];             /* reflectance / transmittance per lane */
    float wo[3], wi[3];           /* local shading frame */
    float u[2];
    uint32_t flags;               /* FL_RENDER_VNDF_PDF */
} fl_bsdf_query;

typedef struct {
    float f_eval[4], pdf_eval;    /* evaluate(wo, wi), pdf(wo, wi) */
    float f_sample[4], pdf_sample;/* sample(wo, u) */
    float wi_sample[3];
    uint32_t sampled_type;        /* BxDFType bits */
} fl_bsdf_result;

int fl_bsdf_query_batch(fl_ctx_t ctx, const fl_bsdf_query *q, fl_bsdf_result *r, uint32_t n);

/* ---- sampler (Sampler::get_1d / get_2d, sampling/sampler.h:14-18) ---------
 * The render path's counter-based sampler, evaluated on the device: for key
 * i = (pixel, sample, bounce) = keys[3i..3i+2] and dimension d < n_dims,
 * out[i * n_dims + d] is the uniform [0, 1) number a render with `seed` draws
 * at that vertex and dimension (a hash of seed, pixel, sample, bounce and
 * dimension; 24-bit mantissa). The reference draws from a seeded mt19937
 * instead (samplers/whitenoise.h:11-29); the two are compared statistically. */
int fl_sampler_draw(fl_ctx_t ctx, uint64_t seed, const uint32_t *keys, uint32_t n, uint32_t n_dims, float *out);

/* The full StandardSurfaceMaterial BSDF (BSDF::evaluate / sample / pdf over
 * its lobe stack, shading/bsdf.cpp:6-151, with flags BSDF_ALL) for n queries
 * of one material. Per query the spectra are given per lane: kd, ks, kt,
 * sheen, coat (from_reflectance_rgb of the material's colours), the forge
 * metal's eta and k, and the lane wavelengths (nm; the thin film's phase). */
typedef struct {
    float kd[4], ks[4], kt[4], sheen[4], coat[4], eta[4], k[4], lam[4];
    float wo[3], wi[3];           /* local shading frame */
    float u[2];
    uint32_t flags;               /* FL_RENDER_VNDF_PDF */
} fl_standard_query;

int fl_bsdf_query_standard(fl_ctx_t ctx, const fl_standard_surface *m, const fl_standard_query *q,
                           fl_bsdf_result *r, uint32_t n);

/* ---- output path (color/spectrumoperations.cpp:58-76,209-286) -------------
 * A pixel's mean XYZ (film sum / spp) goes to a colour space (xyz_to_rgb,
 * :58-66, or xyz_to_aces2065_1, :68-76), through tone_map (:254-286), then
 * gamma_correct (:215-252). The enum values are the reference's. */
enum { FL_COLOR_SRGB = 0, FL_COLOR_ACES2065_1 = 1 };
enum { FL_TONEMAP_NONE = 1, FL_TONEMAP_ACES_BASIC = 2, FL_TONEMAP_REINHARD = 3, FL_TONEMAP_FILMIC = 4 };
enum { FL_GAMMA_NONE = 1, FL_GAMMA_REC_709 = 2, FL_GAMMA_1_8 = 3, FL_GAMMA_2_2 = 4, FL_GAMMA_4_0 = 5 };

typedef struct {
    uint32_t color_space;   /* FL_COLOR_* */
    uint32_t tone_map;      /* FL_TONEMAP_* */
    uint32_t gamma;         /* FL_GAMMA_* */
    uint32_t pad;
} fl_develop_params;

/* Develops the film of item `item` of the context's most recent fl_render /
 * fl_render_batch, still resident on the device: rgb receives width*height*3
 * floats and/or rgba8 width*height*4 bytes (round(clamp(v, 0, 1) * 255),
 * NaN -> 0, alpha 255); either may be NULL. Only the developed pixels cross
 * PCIe (4 B per pixel for rgba8 against the film's 24). */
int fl_develop_film(fl_ctx_t ctx, uint32_t item, const fl_develop_params *params, float *rgb, uint8_t *rgba8);
/* The same for n host pixel means (n*3 floats of XYZ). */
int fl_develop(fl_ctx_t ctx, const float *xyz, uint32_t n, const fl_develop_params *params, float *rgb,
               uint8_t *rgba8);

#ifdef __cplusplus
}
#endif
#endif /* FL_GPU_H */
