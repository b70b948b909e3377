/* fl_scenefile.h -- a scene ingestion wire format for the render path.
 *
 * filianore-core has no scene format of its own: the application that links
 * it builds Scene, materials, illuminants and cameras through the C++ API
 * (core/scene.h, shading/materials, illuminants, cameras/physical.h). This
 * file format carries exactly what fl_scene_create and fl_render take, so a
 * scene assembled in one process (the application, or a converter) can be
 * rendered in another, and renders from the file are bit-identical to renders
 * from the in-memory description.
 *
 * Layout (little endian):
 *   header   8 bytes "FLSCENE\0", uint32 version (1), uint32 section count
 *   section  uint32 tag, uint32 element count, uint64 payload bytes,
 *            then the payload, zero-padded to a multiple of 8 bytes
 * Tags and payloads (the ABI structs of fl_gpu.h, byte for byte):
 *   VERT  float[3] per vertex        INDX  uint32[3] per triangle
 *   TMAT  uint32 per triangle        MATL  fl_material
 *   STDS  fl_standard_surface        ALIT  fl_area_light
 *   DLIT  fl_delta_light             EMIS  float[FL_SPD_N] per spectrum
 *   UVS_  float[2] per vertex        CAMR  fl_camera
 *   TEXR  one texture: uint32 width, uint32 height, width*height*3 bytes
 *         (count = 1; one TEXR section per texture, in index order)
 * Unknown tags are skipped, so later versions can add sections.
 */
#ifndef FL_SCENEFILE_H
#define FL_SCENEFILE_H

#include "fl_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FL_SCENEFILE_VERSION 1

/* Writes desc (and n_cameras cameras, which may be 0) to path. */
int fl_scene_file_write(const char *path, const fl_scene_desc *desc, const fl_camera *cameras, uint32_t n_cameras);

typedef struct fl_scene_file_s *fl_scene_file_t;

/* Reads and checks a scene file: every section's size must match its count,
 * and every index (triangle vertices, materials, lights, emissions, standard
 * surfaces, textures) must be in range. FL_E_INVALID with the reason in
 * fl_scene_file_error otherwise (out is then NULL). */
int fl_scene_file_read(const char *path, fl_scene_file_t *out);
const char *fl_scene_file_error(void);   /* reason of the calling thread's last failed read/write */

/* The file's scene description (pointers into the loaded file, valid until
 * fl_scene_file_free), ready for fl_scene_create. */
const fl_scene_desc *fl_scene_file_desc(fl_scene_file_t f);
/* The file's cameras; returns their number. */
uint32_t fl_scene_file_cameras(fl_scene_file_t f, const fl_camera **cameras);
void fl_scene_file_free(fl_scene_file_t f);

#ifdef __cplusplus
}
#endif

#endif
