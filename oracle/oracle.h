/*
 * oracle.h -- CPU oracle of the render path.  TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker: tests, the smoke check and the CPU-baseline
 * leg of bench.py load it; the product (libfl_gpu.so) never does.  It
 * restates, in plain C++ on the CPU, what the reference computes:
 *   - 60-bin spectra and their tables            color/ (spectrum tables)
 *   - the BxDFs, in fp32 and in fp64              shading/ (BxDFs)
 *   - PathIntegrator::Li with estimate_direct     integrators/
 *   - Embree's role (closest hit) by a binned-SAH BVH over the same
 *     watertight triangle test the GPU uses, plus a brute-force search.
 * The reference's defects are repaired as listed in oracle/README.md.
 */
#ifndef FL_ORACLE_H
#define FL_ORACLE_H

#include <stdint.h>

#include "../include/fl_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene_s *or_scene_t;

/* spectral tables (60 bins, 400..700 nm) */
int or_spectral_bins(void);
void or_cie_bins(float *x, float *y, float *z);                       /* InitializeSpectrals */
void or_refl_basis_bins(int which /* 0..6 WCMYRGB */, float *out);
void or_from_reflectance_rgb(const float rgb[3], float *out60);
void or_from_spd(const float *lam, const float *val, int n, float *out60);
float or_average_spectrum_samples(const float *lam, const float *val, int n, float l0, float l1);
void or_to_xyz(const float *spec60, float xyz[3]);
void or_metal_bins(int metal, float *eta60, float *k60);

/* watertight ray/triangle test (the arithmetic the GPU uses) */
int or_tri_intersect(const float ray[8], const float v0[3], const float v1[3], const float v2[3], float out[3]);

/* scene: triangles + materials + lights (same descriptor as fl_scene_create) */
int or_scene_create(const fl_scene_desc *desc, or_scene_t *out);
void or_scene_destroy(or_scene_t s);
/* rays: n x fl_ray; hits: n x fl_hit.  mode 0 = BVH, 1 = brute force */
int or_intersect(or_scene_t s, const fl_ray *rays, fl_hit *hits, uint32_t n, uint32_t flags, int mode, int threads);
/* best two distinct-triangle hits per ray (for near-tie accounting) */
int or_intersect2(or_scene_t s, const fl_ray *rays, float *t2, uint32_t *id2, uint32_t n, int threads);

/* BxDF queries: precision 32 or 64 */
int or_bsdf_batch(const fl_bsdf_query *q, fl_bsdf_result *r, uint32_t n, int precision);
/* the StandardSurface lobe stack of one material (as fl_bsdf_query_standard) */
int or_bsdf_standard(const fl_standard_surface *m, const fl_standard_query *q, fl_bsdf_result *r, uint32_t n,
                     int precision);

/* Render with the restated PathIntegrator (60-bin spectra, fp32).
 * film_sum/film_sq: W*H*3 floats (CIE XYZ per sample summed); pixels
 * restricted to rows [y0, y1) when y1 > y0. */
typedef struct {
    uint32_t spp;
    int32_t diffuse_depth, reflection_depth, transmission_depth;
    float rr_threshold;
    uint32_t rr_start;
    uint64_t seed;
    int32_t threads;
    uint32_t y0, y1;
    uint32_t x0, x1;
    int32_t integrator;    /* 0 PathIntegrator, 1 NormalIntegrator */
    int32_t mutation;      /* OR_MUT_*: a planted mistake in Li, so that tests can show
                              their checks reject it; 0 = the restatement as written */
} or_render_params;

/* planted mistakes of PathIntegrator::Li (test infrastructure) */
enum {
    OR_MUT_NONE = 0,
    OR_MUT_RR_DIV_Q = 1,           /* roulette survivors divided by q instead of 1 - q */
    OR_MUT_EMIT_ALWAYS = 2,        /* emission added at every hit, not only at bounce 0 / after specular */
    OR_MUT_DEPTH_LATE = 3,         /* the diffuse depth break one bounce late */
    OR_MUT_NO_ETASCALE = 4,        /* etaScale never updated */
    OR_MUT_NO_SPECULAR_EMIT = 5    /* emission only at bounce 0, not after a specular bounce */
};

typedef struct {
    double seconds;
    uint64_t samples, rays;
    int32_t threads;
} or_render_stats;

int or_render(or_scene_t s, const fl_camera *cam, const or_render_params *p, float *film_sum, float *film_sq,
              or_render_stats *st);

/* output path: n pixel means of XYZ to colour space, tone map, gamma
 * (FL_COLOR_*, FL_TONEMAP_*, FL_GAMMA_*); rgba8 as fl_develop */
void or_develop(const float *xyz, uint32_t n, int color_space, int tone_map, int gamma, float *rgb, uint8_t *rgba8);

/* learned proposal (oracle_guide.cpp): parameters are 5 x (64 x 64 + 64)
 * doubles in the GPU's order; records are n x 68 floats (encoded input 64,
 * direction 3, weight 1); bf16 = 1 rounds where the GPU rounds to bf16 */
double or_bf16_round(double x);
void or_guide_encode(const float pos01[3], const float n[3], float lam01, float rough, float *x);
void or_guide_forward(const double *P, const float *x, uint32_t n, int bf16, double *raw);
double or_guide_gradient(const double *P, const float *rec, uint32_t n, int bf16, double *grad);
double or_vmf_mixture_pdf(const double *raw, const double *dir);
double or_vmf_norm(double kappa);
double or_vmf_sample_cos(double kappa, double u);

#ifdef __cplusplus
}
#endif
#endif
