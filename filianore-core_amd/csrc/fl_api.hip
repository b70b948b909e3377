// fl_api.hip -- the C ABI of include/fl_gpu.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "fl_internal.h"

namespace flx {

int set_error(fl_ctx_t ctx, int code, const std::string &msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

int render(fl_scene_s *s, const fl_camera *camera, const fl_render_params *p, fl_film *film, fl_stats *stats);
void release_render_buffers(fl_ctx_t ctx);

} // namespace flx

using flx::set_error;

namespace {

template <typename T>
int upload(fl_ctx_t ctx, T **d, const T *h, size_t n) {
    *d = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc((void **)d, sizeof(T) * n) != hipSuccess) return set_error(ctx, FL_E_OOM, "hipMalloc");
    if (h) {
        hipError_t e = hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice);
        if (e != hipSuccess) return set_error(ctx, FL_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    return FL_OK;
}

void free_scene(fl_scene_s *s) {
    if (!s) return;
    (void)hipFree(s->d_nodes);
    (void)hipFree(s->d_tris);
    (void)hipFree(s->d_verts);
    (void)hipFree(s->d_idx);
    (void)hipFree(s->d_mat);
    (void)hipFree(s->d_light);
    (void)hipFree(s->d_std);
    (void)hipFree(s->d_tex_info);
    (void)hipFree(s->d_tex_data);
    (void)hipFree(s->d_tri_uv);
    (void)hipFree(s->d_texscene);
    (void)hipFree(s->d_mats);
    (void)hipFree(s->d_lights);
    for (auto &t : s->spec) (void)hipFree(t.d_tab);
    delete s;
}

__global__ void __launch_bounds__(256) k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

} // namespace

extern "C" {

int fl_abi_version(void) { return FL_ABI_VERSION; }

int fl_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int fl_probe_copy(fl_ctx_t ctx, uint64_t bytes, uint32_t iters, double *gbps) {
    if (!ctx || !gbps || bytes < 16 || iters == 0) return set_error(ctx, FL_E_INVALID, "fl_probe_copy: bad argument");
    (void)hipSetDevice(ctx->device);
    (void)hipGetLastError();
    const size_t n = (size_t)(bytes / 16);
    uint4 *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, n * 16) != hipSuccess) return set_error(ctx, FL_E_OOM, "fl_probe_copy: source buffer");
    if (hipMalloc(&dst, n * 16) != hipSuccess) {
        (void)hipFree(src);
        return set_error(ctx, FL_E_OOM, "fl_probe_copy: destination buffer");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FL_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = FL_E_HIP;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)ctx->n_cu * 16);
    if (rc == FL_OK) {
        hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, ctx->stream, src, dst, n);   // warm-up
        (void)hipEventRecord(e0, ctx->stream);
        for (uint32_t i = 0; i < iters; ++i)
            hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, ctx->stream, src, dst, n);
        (void)hipEventRecord(e1, ctx->stream);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess || hipGetLastError() != hipSuccess) rc = FL_E_HIP;
    }
    float ms = 0.f;
    if (rc == FL_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.f)
        *gbps = 2.0 * (double)n * 16.0 * iters / (ms * 1e-3) / 1e9;
    else if (rc == FL_OK)
        rc = FL_E_HIP;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    return rc == FL_OK ? FL_OK : set_error(ctx, rc, "fl_probe_copy: HIP failure");
}

int fl_ctx_create(int device, fl_ctx_t *out) {
    if (!out) return FL_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return FL_E_UNSUPPORTED;
    if (device < 0 || device >= n) return FL_E_INVALID;
    if (hipSetDevice(device) != hipSuccess) return FL_E_HIP;
    fl_ctx_s *c = new (std::nothrow) fl_ctx_s();
    if (!c) return FL_E_OOM;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return FL_E_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
    if (c->n_cu <= 0) c->n_cu = 256;
    if (flx::noalloc_scratch_init(c) != FL_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return FL_E_HIP;
    }
    *out = c;
    return FL_OK;
}

void fl_ctx_destroy(fl_ctx_t ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    flx::release_render_buffers(ctx);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->develop_buf) (void)hipFree(ctx->develop_buf);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *fl_last_error(fl_ctx_t ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int fl_scene_create(fl_ctx_t ctx, const fl_scene_desc *desc, fl_scene_t *out) {
    if (!ctx || !desc || !out) return set_error(ctx, FL_E_INVALID, "fl_scene_create: null argument");
    *out = nullptr;
    if (desc->n_triangles && (!desc->vertices || !desc->indices))
        return set_error(ctx, FL_E_INVALID, "fl_scene_create: missing geometry");
    if (desc->n_materials == 0 || !desc->materials) return set_error(ctx, FL_E_INVALID, "fl_scene_create: no materials");
    if (desc->n_lights && (!desc->lights || !desc->emission_spd || desc->n_emission == 0))
        return set_error(ctx, FL_E_INVALID, "fl_scene_create: lights need emission spectra");
    if (desc->n_delta_lights && (!desc->delta_lights || !desc->emission_spd || desc->n_emission == 0))
        return set_error(ctx, FL_E_INVALID, "fl_scene_create: delta lights need emission spectra");
    for (uint32_t l = 0; l < desc->n_delta_lights; ++l) {
        const fl_delta_light &dl = desc->delta_lights[l];
        const float *v = dl.direction;
        const bool has_dir = v[0] != 0.f || v[1] != 0.f || v[2] != 0.f;
        if (dl.type < FL_LIGHT_POINT || dl.type > FL_LIGHT_DIRECTIONAL || dl.emission >= desc->n_emission ||
            dl.shadow_emission > desc->n_emission || (dl.type != FL_LIGHT_POINT && !has_dir))
            return set_error(ctx, FL_E_INVALID, "fl_scene_create: bad delta light");
    }
    for (uint32_t l = 0; l < desc->n_lights; ++l)
        if (desc->lights[l].shadow_emission > desc->n_emission)
            return set_error(ctx, FL_E_INVALID, "fl_scene_create: bad light shadow color");
    if (desc->n_standard_surfaces && !desc->standard_surfaces)
        return set_error(ctx, FL_E_INVALID, "fl_scene_create: missing standard surfaces");
    for (uint32_t m = 0; m < desc->n_materials; ++m)
        if (desc->materials[m].kind == FL_MAT_STANDARD && desc->materials[m].standard >= desc->n_standard_surfaces)
            return set_error(ctx, FL_E_INVALID, "fl_scene_create: standard surface index out of range");
    for (uint32_t j = 0; j < desc->n_standard_surfaces; ++j) {
        const fl_standard_surface &ss = desc->standard_surfaces[j];
        if (ss.fresnel_type < FL_FRESNEL_DIELECTRIC || ss.fresnel_type > FL_FRESNEL_WAVELENGTH_METALLIC ||
            !(ss.ks_ior > 0.f) || (ss.coat_weight > 0.f && !(ss.coat_ior > 0.f)) ||
            (ss.thin_film_thickness > 0.f && !(ss.thin_film_ior > 0.f)))
            return set_error(ctx, FL_E_INVALID, "fl_scene_create: bad standard surface");
    }
    if (desc->bvh_width == 1 || desc->bvh_width > 8)
        return set_error(ctx, FL_E_INVALID, "fl_scene_create: bvh_width must be 0 (8) or 2..8");
    if (desc->n_textures && !desc->textures) return set_error(ctx, FL_E_INVALID, "fl_scene_create: missing textures");
    uint64_t tex_bytes = 0;
    for (uint32_t j = 0; j < desc->n_textures; ++j) {
        const fl_texture &tx = desc->textures[j];
        if (!(tx.rgb && tx.width && tx.height)) continue;
        if (tx.width > 0xffffffffull / 3 / tx.height)   // 3 w h alone beyond 4 GiB (and no wrap below)
            return set_error(ctx, FL_E_INVALID, "fl_scene_create: textures exceed 4 GiB");
        tex_bytes += (3ull * tx.width * tx.height + 3) & ~3ull;
        if (tex_bytes > 0xffffffffull) break;
    }
    if (tex_bytes > 0xffffffffull) return set_error(ctx, FL_E_INVALID, "fl_scene_create: textures exceed 4 GiB");
    for (uint32_t m = 0; m < desc->n_materials; ++m)
        if (desc->materials[m].spd_r > desc->n_emission || desc->materials[m].spd_t > desc->n_emission)
            return set_error(ctx, FL_E_INVALID, "fl_scene_create: material spectrum index out of range");
    for (uint32_t m = 0; m < desc->n_materials; ++m)
        if (desc->materials[m].kind == FL_MAT_LAMBERT && desc->materials[m].texture > desc->n_textures)
            return set_error(ctx, FL_E_INVALID, "fl_scene_create: texture index out of range");
    for (uint32_t j = 0; j < desc->n_standard_surfaces; ++j) {
        const fl_standard_surface &ss = desc->standard_surfaces[j];
        const uint32_t t[5] = {ss.kd_texture, ss.ks_texture, ss.sheen_texture, ss.kt_texture, ss.coat_texture};
        for (uint32_t k = 0; k < 5; ++k)
            if (t[k] > desc->n_textures)
                return set_error(ctx, FL_E_INVALID, "fl_scene_create: texture index out of range");
    }
    for (uint64_t t = 0; t < 3ull * desc->n_triangles; ++t)
        if (desc->indices[t] >= desc->n_vertices) return set_error(ctx, FL_E_INVALID, "fl_scene_create: index out of range");
    if (desc->tri_material)
        for (uint32_t t = 0; t < desc->n_triangles; ++t)
            if (desc->tri_material[t] >= desc->n_materials && desc->tri_material[t] != fl::kNoMaterial)
                return set_error(ctx, FL_E_INVALID, "fl_scene_create: material index out of range");
    (void)hipSetDevice(ctx->device);
    fl_scene_s *s = new (std::nothrow) fl_scene_s();
    if (!s) return set_error(ctx, FL_E_OOM, "scene");
    s->ctx = ctx;
    s->n_tris = desc->n_triangles;
    s->n_verts = desc->n_vertices;
    s->n_mats = desc->n_materials;
    s->n_lights = desc->n_lights + desc->n_delta_lights;

    // per-triangle light index + 1, lights' geometry (Triangle::sample / area)
    std::vector<uint32_t> tri_light(desc->n_triangles, 0u);
    s->h_lights.assign(s->n_lights, DevLight{});
    for (uint32_t l = 0; l < desc->n_lights; ++l) {
        const fl_area_light &al = desc->lights[l];
        if (al.tri >= desc->n_triangles || al.emission >= desc->n_emission) {
            free_scene(s);
            return set_error(ctx, FL_E_INVALID, "fl_scene_create: bad light");
        }
        tri_light[al.tri] = l + 1;
        DevLight &d = s->h_lights[l];
        const float *P[3];
        for (int k = 0; k < 3; ++k) P[k] = desc->vertices + 3ull * desc->indices[3ull * al.tri + k];
        for (int k = 0; k < 3; ++k) { d.p0[k] = P[0][k]; d.p1[k] = P[1][k]; d.p2[k] = P[2][k]; }
        // cross evaluated in double then rounded, as maths/geometry.h:213-219 does
        double e1[3], e2[3];
        for (int k = 0; k < 3; ++k) { e1[k] = (double)(float)(P[1][k] - P[0][k]); e2[k] = (double)(float)(P[2][k] - P[0][k]); }
        float c[3] = {(float)(e1[1] * e2[2] - e1[2] * e2[1]), (float)(e1[2] * e2[0] - e1[0] * e2[2]),
                      (float)(e1[0] * e2[1] - e1[1] * e2[0])};
        float len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (int k = 0; k < 3; ++k) d.n[k] = c[k] / len;
        d.area = 0.5f * len;
        d.intensity = al.intensity;
        d.tri = al.tri;
        d.em_slot = 0; // set after tables are sized
    }
    std::vector<DevMaterial> mats;
    s->h_std.resize(desc->n_standard_surfaces);
    for (uint32_t m = 0; m < desc->n_materials; ++m) s->has_standard = s->has_standard || desc->materials[m].kind == FL_MAT_STANDARD;
    for (uint32_t j = 0; j < desc->n_standard_surfaces; ++j) s->h_std[j] = flx::standard_params(desc->standard_surfaces[j]);
    for (uint32_t m = 0; m < desc->n_materials; ++m) {
        const fl_material &fm = desc->materials[m];
        if (fm.kind == FL_MAT_LAMBERT && fm.texture) s->has_textures = true;
        if (fm.kind == FL_MAT_STANDARD)
            for (uint32_t k = 0; k < (uint32_t)fl::SS_ETA; ++k)
                if (s->h_std[fm.standard].tex[k]) s->has_textures = true;
    }
    flx::build_spectral_tables(s, desc, FL_SPECTRAL_PARITY, mats);
    flx::build_spectral_tables(s, desc, FL_SPECTRAL_HERO, mats);
    s->h_mats = mats;
    const uint32_t em0 = 3 + 2 * desc->n_materials;   // table slot of emission 0
    for (uint32_t l = 0; l < desc->n_lights; ++l) {
        DevLight &d = s->h_lights[l];
        d.em_slot = em0 + desc->lights[l].emission;
        d.shadow_slot = desc->lights[l].shadow_emission ? em0 + desc->lights[l].shadow_emission - 1 : 0u;
    }
    // delta lights: point.cpp:6-10, spot.cpp:6-21, directional.cpp:6-13
    for (uint32_t l = 0; l < desc->n_delta_lights; ++l) {
        const fl_delta_light &dl = desc->delta_lights[l];
        DevLight &d = s->h_lights[desc->n_lights + l];
        d.type = (uint32_t)dl.type;
        d.tri = fl::kNoHit;
        d.intensity = dl.intensity;
        d.em_slot = em0 + dl.emission;
        d.shadow_slot = dl.shadow_emission ? em0 + dl.shadow_emission - 1 : 0u;
        if (dl.type == FL_LIGHT_DIRECTIONAL) {
            // dirIllum = normalize(-vector_transform(dir)): toward the light
            const float v[3] = {-dl.direction[0], -dl.direction[1], -dl.direction[2]};
            const float len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            for (int k = 0; k < 3; ++k) d.p0[k] = v[k] / len;
        } else {
            for (int k = 0; k < 3; ++k) { d.p0[k] = dl.position[k]; d.p1[k] = dl.direction[k]; }
            d.cos_cone = dl.cos_cone;
            d.cos_penumbra = dl.cos_penumbra;
        }
    }
    for (const DevLight &d : s->h_lights) s->shadow_colors = s->shadow_colors || d.shadow_slot != 0u;

    int rc;
    std::vector<uint32_t> mat_h;
    // the build inputs stay on the device for rebuilds at another node width
    s->bvh_width = desc->bvh_width ? desc->bvh_width : 8u;
    if ((rc = upload(ctx, &s->d_verts, desc->vertices, 3ull * desc->n_vertices)) ||
        (rc = upload(ctx, &s->d_idx, desc->indices, 3ull * desc->n_triangles)) ||
        (rc = upload(ctx, &s->d_light, tri_light.data(), desc->n_triangles))) {
        free_scene(s);
        return rc;
    }
    if (desc->tri_material) {
        // a material without BSDF (FL_MAT_NONE) is a pass-through primitive
        mat_h.assign(desc->tri_material, desc->tri_material + desc->n_triangles);
        for (auto &m : mat_h)
            if (m != fl::kNoMaterial && desc->materials[m].kind == FL_MAT_NONE) m = fl::kNoMaterial;
        if ((rc = upload(ctx, &s->d_mat, mat_h.data(), desc->n_triangles))) {
            free_scene(s);
            return rc;
        }
    }
    rc = flx::build_bvh(ctx, s, s->d_verts, s->d_idx, s->d_mat, s->d_light);
    if (rc) { free_scene(s); return rc; }
    if (s->has_textures) {
        // texels of every texture, each padded to 4 bytes; per-triangle uvs
        std::vector<fl::DevTexture> info(desc->n_textures);
        std::vector<uint8_t> data((size_t)tex_bytes);
        uint64_t off = 0;
        for (uint32_t j = 0; j < desc->n_textures; ++j) {
            const fl_texture &tx = desc->textures[j];
            info[j] = fl::DevTexture{(uint32_t)off, 0u, 0u, 0u};
            if (!(tx.rgb && tx.width && tx.height)) continue;   // black
            const size_t nb = 3ull * tx.width * tx.height;
            memcpy(data.data() + off, tx.rgb, nb);
            info[j].width = tx.width;
            info[j].height = tx.height;
            off += (nb + 3) & ~(size_t)3;
        }
        std::vector<float2> tri_uv;
        if (desc->uvs) {
            tri_uv.resize(3ull * desc->n_triangles);
            for (size_t t = 0; t < desc->n_triangles; ++t)
                for (int k = 0; k < 3; ++k) {
                    const size_t vtx = desc->indices[3 * t + k];
                    tri_uv[3 * t + k] = make_float2(desc->uvs[2 * vtx], desc->uvs[2 * vtx + 1]);
                }
        }
        if (data.empty()) data.resize(4);
        if ((rc = upload(ctx, &s->d_tex_info, info.data(), info.size())) ||
            (rc = upload(ctx, &s->d_tex_data, data.data(), data.size())) ||
            (!tri_uv.empty() && (rc = upload(ctx, &s->d_tri_uv, tri_uv.data(), tri_uv.size())))) {
            free_scene(s);
            return rc;
        }
        fl::TexScene ts{s->d_tex_info, s->d_tex_data, s->d_tri_uv};
        if ((rc = upload(ctx, &s->d_texscene, &ts, 1))) {
            free_scene(s);
            return rc;
        }
    }
    if ((rc = upload(ctx, &s->d_mats, s->h_mats.data(), s->h_mats.size())) ||
        (rc = upload(ctx, &s->d_lights, s->h_lights.data(), s->h_lights.size())) ||
        (rc = upload(ctx, &s->d_std, s->h_std.data(), s->h_std.size())) ||
        (rc = upload(ctx, &s->spec[0].d_tab, s->spec[0].host.data(), s->spec[0].host.size())) ||
        (rc = upload(ctx, &s->spec[1].d_tab, s->spec[1].host.data(), s->spec[1].host.size()))) {
        free_scene(s);
        return rc;
    }
    for (int k = 0; k < 3; ++k) { s->bmin[k] = fl::kInf; s->bmax[k] = -fl::kInf; }
    for (uint32_t v = 0; v < desc->n_vertices; ++v)
        for (int k = 0; k < 3; ++k) {
            s->bmin[k] = std::fmin(s->bmin[k], desc->vertices[3ull * v + k]);
            s->bmax[k] = std::fmax(s->bmax[k], desc->vertices[3ull * v + k]);
        }
    *out = s;
    return FL_OK;
}

void fl_scene_destroy(fl_scene_t s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    free_scene(s);
}

int fl_scene_get_info(fl_scene_t s, fl_scene_info *out) {
    if (!s || !out) return FL_E_INVALID;
    memset(out, 0, sizeof(*out));
    out->n_triangles = s->n_tris;
    out->n_nodes = s->n_nodes;
    out->n_leaf_tris = s->n_tris;
    out->depth = s->depth;
    out->node_bytes = (uint64_t)s->n_nodes * sizeof(fl::Node8);
    out->tri_bytes = (uint64_t)s->n_tris * sizeof(fl::TriRec);
    out->build_ms = s->build_ms;
    for (int k = 0; k < 3; ++k) { out->bounds_min[k] = s->bmin[k]; out->bounds_max[k] = s->bmax[k]; }
    return FL_OK;
}

int fl_scene_download_bvh(fl_scene_t s, void *nodes, size_t node_bytes, void *tris, size_t tri_bytes) {
    if (!s) return FL_E_INVALID;
    fl_ctx_t ctx = s->ctx;
    size_t nb = (size_t)s->n_nodes * sizeof(fl::Node8), tb = (size_t)s->n_tris * sizeof(fl::TriRec);
    if ((nodes && node_bytes < nb) || (tris && tri_bytes < tb)) return set_error(ctx, FL_E_INVALID, "buffer too small");
    (void)hipSetDevice(ctx->device);
    if (nodes) FL_HIP_CHECK(ctx, hipMemcpy(nodes, s->d_nodes, nb, hipMemcpyDeviceToHost));
    if (tris && tb) FL_HIP_CHECK(ctx, hipMemcpy(tris, s->d_tris, tb, hipMemcpyDeviceToHost));
    return FL_OK;
}

int fl_intersect_device(fl_scene_t s, const fl_ray *d_rays, fl_hit *d_hits, uint32_t n, uint32_t flags) {
    if (!s || (n && (!d_rays || !d_hits))) return FL_E_INVALID;
    (void)hipSetDevice(s->ctx->device);
    return flx::trace_rays(s->ctx, s, d_rays, d_hits, n, flags & ~FL_COUNTERS, nullptr);
}

int fl_intersect_device_noalloc(fl_scene_t s, const fl_ray *d_rays, fl_hit *d_hits, uint32_t n, uint32_t flags,
                                void *stream) {
    if (!s || (n && (!d_rays || !d_hits))) return FL_E_INVALID;
    if (flags & FL_COUNTERS)
        return set_error(s->ctx, FL_E_INVALID, "fl_intersect_device_noalloc: FL_COUNTERS needs fl_intersect");
    (void)hipSetDevice(s->ctx->device);
    return flx::trace_rays(s->ctx, s, d_rays, d_hits, n, flags, nullptr, true, static_cast<hipStream_t>(stream));
}

int fl_intersect(fl_scene_t s, const fl_ray *rays, fl_hit *hits, uint32_t n, uint32_t flags, uint64_t *counters) {
    if (!s || (n && (!rays || !hits))) return FL_E_INVALID;
    fl_ctx_t ctx = s->ctx;
    if (n == 0) return FL_OK;
    (void)hipSetDevice(ctx->device);
    fl_ray *dr = nullptr;
    fl_hit *dh = nullptr;
    unsigned long long *dc = nullptr;
    int rc = FL_OK;
    if (hipMalloc((void **)&dr, sizeof(fl_ray) * (size_t)n) != hipSuccess ||
        hipMalloc((void **)&dh, sizeof(fl_hit) * (size_t)n) != hipSuccess ||
        hipMalloc((void **)&dc, 64) != hipSuccess) {
        rc = set_error(ctx, FL_E_OOM, "fl_intersect buffers");
    }
    if (!rc && hipMemcpy(dr, rays, sizeof(fl_ray) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        rc = set_error(ctx, FL_E_HIP, "upload rays");
    if (!rc && hipMemset(dc, 0, 64) != hipSuccess) rc = set_error(ctx, FL_E_HIP, "counters");
    if (!rc) rc = flx::trace_rays(ctx, s, dr, dh, n, flags, (flags & FL_COUNTERS) ? dc : nullptr);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = set_error(ctx, FL_E_HIP, "trace");
    if (!rc && hipMemcpy(hits, dh, sizeof(fl_hit) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_error(ctx, FL_E_HIP, "download hits");
    if (!rc && counters) {
        unsigned long long c[3];
        if (hipMemcpy(c, dc, sizeof(c), hipMemcpyDeviceToHost) == hipSuccess) {
            counters[0] = c[0];
            counters[1] = c[1];
        }
    }
    (void)hipFree(dr);
    (void)hipFree(dh);
    (void)hipFree(dc);
    return rc;
}

int fl_render(fl_scene_t s, const fl_camera *camera, const fl_render_params *params, fl_film *film, fl_stats *stats) {
    if (!s) return FL_E_INVALID;
    (void)hipSetDevice(s->ctx->device);
    return flx::render(s, camera, params, film, stats);
}

int fl_bsdf_query_standard(fl_ctx_t ctx, const fl_standard_surface *m, const fl_standard_query *q,
                           fl_bsdf_result *r, uint32_t n) {
    if (!ctx || !m || (n && (!q || !r))) return FL_E_INVALID;
    if (m->fresnel_type < FL_FRESNEL_DIELECTRIC || m->fresnel_type > FL_FRESNEL_WAVELENGTH_METALLIC)
        return set_error(ctx, FL_E_INVALID, "fl_bsdf_query_standard: bad fresnel type");
    if (n == 0) return FL_OK;
    (void)hipSetDevice(ctx->device);
    fl_standard_query *dq = nullptr;
    fl_bsdf_result *dr = nullptr;
    int rc = FL_OK;
    if (hipMalloc((void **)&dq, sizeof(*q) * (size_t)n) != hipSuccess ||
        hipMalloc((void **)&dr, sizeof(*r) * (size_t)n) != hipSuccess)
        rc = set_error(ctx, FL_E_OOM, "bsdf buffers");
    if (!rc && hipMemcpy(dq, q, sizeof(*q) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        rc = set_error(ctx, FL_E_HIP, "upload");
    if (!rc) rc = flx::bsdf_queries_standard(ctx, m, dq, dr, n);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = set_error(ctx, FL_E_HIP, "bsdf kernel");
    if (!rc && hipMemcpy(r, dr, sizeof(*r) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_error(ctx, FL_E_HIP, "download");
    (void)hipFree(dq);
    (void)hipFree(dr);
    return rc;
}

int fl_bsdf_query_batch(fl_ctx_t ctx, const fl_bsdf_query *q, fl_bsdf_result *r, uint32_t n) {
    if (!ctx || (n && (!q || !r))) return FL_E_INVALID;
    if (n == 0) return FL_OK;
    (void)hipSetDevice(ctx->device);
    fl_bsdf_query *dq = nullptr;
    fl_bsdf_result *dr = nullptr;
    int rc = FL_OK;
    if (hipMalloc((void **)&dq, sizeof(*q) * (size_t)n) != hipSuccess ||
        hipMalloc((void **)&dr, sizeof(*r) * (size_t)n) != hipSuccess)
        rc = set_error(ctx, FL_E_OOM, "bsdf buffers");
    if (!rc && hipMemcpy(dq, q, sizeof(*q) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        rc = set_error(ctx, FL_E_HIP, "upload");
    if (!rc) rc = flx::bsdf_queries(ctx, dq, dr, n);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = set_error(ctx, FL_E_HIP, "bsdf kernel");
    if (!rc && hipMemcpy(r, dr, sizeof(*r) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_error(ctx, FL_E_HIP, "download");
    (void)hipFree(dq);
    (void)hipFree(dr);
    return rc;
}

int fl_sampler_draw(fl_ctx_t ctx, uint64_t seed, const uint32_t *keys, uint32_t n, uint32_t n_dims, float *out) {
    if (!ctx) return FL_E_INVALID;
    if ((n && (!keys || !out)) || n_dims == 0 || n_dims > 64)
        return set_error(ctx, FL_E_INVALID, "fl_sampler_draw: null arrays or n_dims outside 1..64");
    if (n == 0) return FL_OK;
    (void)hipSetDevice(ctx->device);
    uint32_t *dk = nullptr;
    float *dout = nullptr;
    const size_t nk = 3ull * n, no = (size_t)n * n_dims;
    int rc = FL_OK;
    if (hipMalloc((void **)&dk, nk * 4) != hipSuccess || hipMalloc((void **)&dout, no * 4) != hipSuccess)
        rc = set_error(ctx, FL_E_OOM, "sampler buffers");
    if (!rc && hipMemcpy(dk, keys, nk * 4, hipMemcpyHostToDevice) != hipSuccess) rc = set_error(ctx, FL_E_HIP, "upload");
    if (!rc) rc = flx::sampler_draw(ctx, seed, dk, n, n_dims, dout);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = set_error(ctx, FL_E_HIP, "sampler kernel");
    if (!rc && hipMemcpy(out, dout, no * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = set_error(ctx, FL_E_HIP, "download");
    (void)hipFree(dk);
    (void)hipFree(dout);
    return rc;
}

int fl_render_batch(fl_ctx_t ctx, const fl_batch_item *items, uint32_t n, const fl_render_params *params,
                    fl_stats *stats) {
    if (!ctx) return FL_E_INVALID;
    if (!items || n == 0 || !params) return set_error(ctx, FL_E_INVALID, "fl_render_batch: no items or params");
    std::vector<flx::HostItem> hi(n);
    for (uint32_t k = 0; k < n; ++k)
        hi[k] = flx::HostItem{items[k].scene, items[k].camera, items[k].film, items[k].spp ? items[k].spp : params->spp};
    (void)hipSetDevice(ctx->device);
    return flx::render_items(ctx, hi.data(), n, params, stats);
}
int fl_autotune(fl_scene_t s, const fl_camera *camera, const fl_render_params *params, fl_tune_result *out) {
    if (!s) return FL_E_INVALID;
    if (!camera || !params) return set_error(s->ctx, FL_E_INVALID, "fl_autotune: null argument");
    (void)hipSetDevice(s->ctx->device);
    flx::HostItem h{s, *camera, nullptr, params->spp};
    return flx::autotune_items(s->ctx, &h, 1, params, out);
}

int fl_autotune_batch(fl_ctx_t ctx, const fl_batch_item *items, uint32_t n, const fl_render_params *params,
                      fl_tune_result *out) {
    if (!ctx) return FL_E_INVALID;
    if (!items || n == 0 || !params) return set_error(ctx, FL_E_INVALID, "fl_autotune_batch: no items or params");
    std::vector<flx::HostItem> hi(n);
    for (uint32_t k = 0; k < n; ++k)
        hi[k] = flx::HostItem{items[k].scene, items[k].camera, nullptr, items[k].spp ? items[k].spp : params->spp};
    (void)hipSetDevice(ctx->device);
    return flx::autotune_items(ctx, hi.data(), n, params, out);
}

} // extern "C"
