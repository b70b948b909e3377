/*
 * fl_scenegen.h -- procedural scenes of the render-path workloads.
 *
 * The reference has no scene files; its only geometry helper is
 * ShapeCreator::create_quad (creators/shapecreator.cpp:7-38: two triangles
 * over [-1,1]^2, local normal -z), which these generators use for walls and
 * lights.  One generator output feeds both the GPU path (fl_scene_create) and
 * the CPU oracle, so both see identical float buffers.
 */
#ifndef FL_SCENEGEN_H
#define FL_SCENEGEN_H

#include "fl_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FL_GEN_CORNELL = 0,   /* small Cornell box: 36 tris, 256^2, 16 spp, depth 8 */
    FL_GEN_SPHERES = 1,   /* 8 displaced icospheres (subdiv 7) + ground: 2,621,442 tris */
    FL_GEN_BATCH = 2,     /* one scene of the 64-scene batch (index = opt.scene_index) */
    FL_GEN_GUIDED = 3,    /* spheres inside a closed room lit through a ceiling slot from an attic */
    FL_GEN_TERRAIN = 4,   /* fBm heightfield + 200 dispersive prisms + metals */
    FL_GEN_GUIDED_SKYLIGHT = 5 /* spheres in a closed 12-tri room under a 5% ceiling strip, light above it */
};

typedef struct {
    int32_t subdiv;        /* icosphere subdivision (-1: config default) */
    int32_t n_spheres;     /* (-1: config default) */
    int32_t grid_x, grid_y;/* terrain grid (0: default 4096 x 2048) */
    uint32_t width, height;/* camera resolution (0: config default) */
    uint32_t scene_index;  /* batch scene index 0..63 */
    uint32_t camera_index; /* terrain: camera 0..7 */
    uint64_t seed;         /* base seed (0: the config's fixed seeds) */
} fl_gen_options;

typedef struct {
    float *vertices;
    uint32_t n_vertices;
    uint32_t *indices;
    uint32_t n_triangles;
    uint32_t *tri_material;
    fl_material *materials;
    uint32_t n_materials;
    fl_area_light *lights;
    uint32_t n_lights;
    float *emission_spd;
    uint32_t n_emission;
    fl_camera camera;
    uint32_t spp;          /* config default samples per pixel */
    int32_t depth;         /* config default max depth (all three RenderParams depths) */
} fl_gen_scene;

void fl_gen_default_options(fl_gen_options *opt);
int fl_gen_config(int config, const fl_gen_options *opt, fl_gen_scene *out);
void fl_gen_free(fl_gen_scene *s);
/* Fills an fl_scene_desc that points into s (no copies). */
void fl_gen_describe(const fl_gen_scene *s, fl_scene_desc *desc);
/* Planck blackbody on 360..830 nm @1 nm, normalised to peak 1. */
void fl_gen_blackbody(float temperature_k, float *spd /* FL_SPD_N */);

#ifdef __cplusplus
}
#endif
#endif
