// fl_scenefile.cpp -- the scene wire format of include/fl_scenefile.h (host code).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fl_scenefile.h"

namespace {

thread_local std::string g_error;

int fail(const std::string &msg) {
    g_error = msg;
    return FL_E_INVALID;
}

constexpr char kMagic[8] = {'F', 'L', 'S', 'C', 'E', 'N', 'E', '\0'};

constexpr uint32_t tag(const char (&s)[5]) {
    return (uint32_t)(uint8_t)s[0] | ((uint32_t)(uint8_t)s[1] << 8) | ((uint32_t)(uint8_t)s[2] << 16) |
           ((uint32_t)(uint8_t)s[3] << 24);
}

constexpr uint32_t T_VERT = tag("VERT"), T_INDX = tag("INDX"), T_TMAT = tag("TMAT"), T_MATL = tag("MATL"),
                   T_STDS = tag("STDS"), T_ALIT = tag("ALIT"), T_DLIT = tag("DLIT"), T_EMIS = tag("EMIS"),
                   T_UVS = tag("UVS_"), T_CAMR = tag("CAMR"), T_TEXR = tag("TEXR");

struct SectionHeader {
    uint32_t tag, count;
    uint64_t bytes;
};

class Writer {
public:
    explicit Writer(FILE *f) : f_(f) {}
    bool section(uint32_t t, uint32_t count, const void *a, size_t na, const void *b = nullptr, size_t nb = 0) {
        SectionHeader h{t, count, (uint64_t)(na + nb)};
        static const uint8_t zero[8] = {};
        const size_t pad = (8 - (na + nb) % 8) % 8;
        return fwrite(&h, sizeof(h), 1, f_) == 1 && (na == 0 || fwrite(a, 1, na, f_) == na) &&
               (nb == 0 || fwrite(b, 1, nb, f_) == nb) && (pad == 0 || fwrite(zero, 1, pad, f_) == pad);
    }

private:
    FILE *f_;
};

} // namespace

struct fl_scene_file_s {
    std::vector<uint8_t> bytes;         // the whole file; the desc points into it
    std::vector<fl_texture> textures;
    fl_scene_desc desc{};
    const fl_camera *cameras = nullptr;
    uint32_t n_cameras = 0;
};

extern "C" {

const char *fl_scene_file_error(void) { return g_error.c_str(); }

int fl_scene_file_write(const char *path, const fl_scene_desc *d, const fl_camera *cameras, uint32_t n_cameras) {
    if (!path || !d || (n_cameras && !cameras)) return fail("fl_scene_file_write: null argument");
    FILE *f = fopen(path, "wb");
    if (!f) return fail(std::string("fl_scene_file_write: cannot open ") + path);
    uint32_t n_sections = 0;
    const uint32_t version = FL_SCENEFILE_VERSION;
    bool ok = fwrite(kMagic, 1, 8, f) == 8 && fwrite(&version, 4, 1, f) == 1 && fwrite(&n_sections, 4, 1, f) == 1;
    Writer w(f);
    auto put = [&](uint32_t t, uint32_t count, const void *p, size_t elem) {
        if (!ok || count == 0) return;
        ok = p && w.section(t, count, p, (size_t)count * elem);
        ++n_sections;
    };
    put(T_VERT, d->n_vertices, d->vertices, 3 * sizeof(float));
    put(T_INDX, d->n_triangles, d->indices, 3 * sizeof(uint32_t));
    if (d->tri_material) put(T_TMAT, d->n_triangles, d->tri_material, sizeof(uint32_t));
    put(T_MATL, d->n_materials, d->materials, sizeof(fl_material));
    put(T_STDS, d->n_standard_surfaces, d->standard_surfaces, sizeof(fl_standard_surface));
    put(T_ALIT, d->n_lights, d->lights, sizeof(fl_area_light));
    put(T_DLIT, d->n_delta_lights, d->delta_lights, sizeof(fl_delta_light));
    put(T_EMIS, d->n_emission, d->emission_spd, FL_SPD_N * sizeof(float));
    if (d->uvs) put(T_UVS, d->n_vertices, d->uvs, 2 * sizeof(float));
    put(T_CAMR, n_cameras, cameras, sizeof(fl_camera));
    for (uint32_t j = 0; ok && j < d->n_textures; ++j) {
        if (!d->textures) {
            ok = false;
            break;
        }
        const fl_texture &t = d->textures[j];
        const bool has = t.rgb && t.width && t.height;
        const uint32_t wh[2] = {has ? t.width : 0u, has ? t.height : 0u};
        ok = w.section(T_TEXR, 1, wh, sizeof(wh), t.rgb, has ? 3ull * t.width * t.height : 0);
        ++n_sections;
    }
    // the section count sits after the magic and version
    ok = ok && fseek(f, 12, SEEK_SET) == 0 && fwrite(&n_sections, 4, 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(std::string("fl_scene_file_write: write failed: ") + path);
    return FL_OK;
}

int fl_scene_file_read(const char *path, fl_scene_file_t *out) {
    if (!path || !out) return fail("fl_scene_file_read: null argument");
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) return fail(std::string("fl_scene_file_read: cannot open ") + path);
    fl_scene_file_s *s = new fl_scene_file_s();
    if (fseek(f, 0, SEEK_END) == 0) {
        const long n = ftell(f);
        if (n > 0) {
            s->bytes.resize((size_t)n);
            rewind(f);
            if (fread(s->bytes.data(), 1, (size_t)n, f) != (size_t)n) s->bytes.clear();
        }
    }
    fclose(f);
    auto bad = [&](const std::string &m) {
        delete s;
        return fail("fl_scene_file_read: " + m);
    };
    const std::vector<uint8_t> &B = s->bytes;
    if (B.size() < 16 || memcmp(B.data(), kMagic, 8) != 0) return bad("not a scene file");
    uint32_t version, n_sections;
    memcpy(&version, B.data() + 8, 4);
    memcpy(&n_sections, B.data() + 12, 4);
    if (version != FL_SCENEFILE_VERSION) return bad("unsupported version " + std::to_string(version));
    fl_scene_desc &d = s->desc;
    size_t off = 16;
    // every tag but TEXR appears at most once; the per-triangle and
    // per-vertex sections must match the INDX and VERT counts
    std::vector<uint32_t> seen;
    uint32_t n_tmat = 0, n_uvs = 0;
    for (uint32_t k = 0; k < n_sections; ++k) {
        SectionHeader h;
        if (B.size() - off < sizeof(h)) return bad("truncated section header");
        memcpy(&h, B.data() + off, sizeof(h));
        off += sizeof(h);
        if (h.bytes > B.size() - off) return bad("truncated section");
        const uint8_t *p = B.data() + off;
        if (h.tag != T_TEXR) {
            if (std::find(seen.begin(), seen.end(), h.tag) != seen.end()) return bad("duplicate section");
            seen.push_back(h.tag);
        }
        auto expect = [&](size_t elem) { return (uint64_t)h.count * elem == h.bytes; };
        // payloads start 8-byte aligned (header 16, sections padded to 8), so
        // the casts below are aligned for every struct of fl_gpu.h
        bool size_ok = true;
        if (h.tag == T_VERT) { size_ok = expect(12); d.vertices = (const float *)p; d.n_vertices = h.count; }
        else if (h.tag == T_INDX) { size_ok = expect(12); d.indices = (const uint32_t *)p; d.n_triangles = h.count; }
        else if (h.tag == T_TMAT) { size_ok = expect(4); d.tri_material = (const uint32_t *)p; n_tmat = h.count; }
        else if (h.tag == T_MATL) { size_ok = expect(sizeof(fl_material)); d.materials = (const fl_material *)p; d.n_materials = h.count; }
        else if (h.tag == T_STDS) {
            size_ok = expect(sizeof(fl_standard_surface));
            d.standard_surfaces = (const fl_standard_surface *)p;
            d.n_standard_surfaces = h.count;
        }
        else if (h.tag == T_ALIT) { size_ok = expect(sizeof(fl_area_light)); d.lights = (const fl_area_light *)p; d.n_lights = h.count; }
        else if (h.tag == T_DLIT) { size_ok = expect(sizeof(fl_delta_light)); d.delta_lights = (const fl_delta_light *)p; d.n_delta_lights = h.count; }
        else if (h.tag == T_EMIS) { size_ok = expect(FL_SPD_N * 4); d.emission_spd = (const float *)p; d.n_emission = h.count; }
        else if (h.tag == T_UVS) { size_ok = expect(8); d.uvs = (const float *)p; n_uvs = h.count; }
        else if (h.tag == T_CAMR) { size_ok = expect(sizeof(fl_camera)); s->cameras = (const fl_camera *)p; s->n_cameras = h.count; }
        else if (h.tag == T_TEXR) {
            uint32_t wh[2] = {0, 0};
            size_ok = h.count == 1 && h.bytes >= 8;
            if (size_ok) {
                memcpy(wh, p, 8);
                // 3 w h bytes, computed only where it cannot wrap
                size_ok = wh[1] == 0 || wh[0] <= (UINT64_MAX - 8) / 3 / wh[1];
                size_ok = size_ok && h.bytes == 8 + 3ull * wh[0] * wh[1];
            }
            fl_texture t{};
            t.width = wh[0];
            t.height = wh[1];
            t.rgb = (wh[0] && wh[1]) ? p + 8 : nullptr;
            s->textures.push_back(t);
        }
        if (!size_ok) return bad("section size does not match its count");
        off += (size_t)h.bytes;
        off += (8 - off % 8) % 8;
        if (off > B.size()) return bad("truncated padding");
    }
    d.textures = s->textures.empty() ? nullptr : s->textures.data();
    d.n_textures = (uint32_t)s->textures.size();
    // every reference in range (fl_scene_create checks them again)
    if (d.n_triangles && (!d.vertices || !d.indices)) return bad("triangles without vertices");
    if (d.tri_material && n_tmat != d.n_triangles) return bad("material section does not match the triangle count");
    if (d.uvs && n_uvs != d.n_vertices) return bad("uv section does not match the vertex count");
    if (!d.materials || d.n_materials == 0) return bad("no materials");
    for (uint64_t i = 0; i < 3ull * d.n_triangles; ++i)
        if (d.indices[i] >= d.n_vertices) return bad("vertex index out of range");
    if (d.tri_material)
        for (uint32_t t = 0; t < d.n_triangles; ++t)
            if (d.tri_material[t] >= d.n_materials && d.tri_material[t] != 0xffffu) return bad("material index out of range");
    for (uint32_t m = 0; m < d.n_materials; ++m) {
        const fl_material &fm = d.materials[m];
        if (fm.kind == FL_MAT_STANDARD && fm.standard >= d.n_standard_surfaces) return bad("standard surface index out of range");
        if (fm.kind == FL_MAT_LAMBERT && fm.texture > d.n_textures) return bad("texture index out of range");
        if (fm.spd_r > d.n_emission || fm.spd_t > d.n_emission) return bad("material spectrum index out of range");
    }
    for (uint32_t j = 0; j < d.n_standard_surfaces; ++j) {
        const fl_standard_surface &ss = d.standard_surfaces[j];
        for (uint32_t t : {ss.kd_texture, ss.ks_texture, ss.sheen_texture, ss.kt_texture, ss.coat_texture})
            if (t > d.n_textures) return bad("texture index out of range");
    }
    for (uint32_t l = 0; l < d.n_lights; ++l)
        if (d.lights[l].tri >= d.n_triangles || d.lights[l].emission >= d.n_emission ||
            d.lights[l].shadow_emission > d.n_emission)
            return bad("area light reference out of range");
    for (uint32_t l = 0; l < d.n_delta_lights; ++l)
        if (d.delta_lights[l].emission >= d.n_emission || d.delta_lights[l].shadow_emission > d.n_emission)
            return bad("delta light emission out of range");
    *out = s;
    return FL_OK;
}

const fl_scene_desc *fl_scene_file_desc(fl_scene_file_t f) { return f ? &f->desc : nullptr; }

uint32_t fl_scene_file_cameras(fl_scene_file_t f, const fl_camera **cameras) {
    if (cameras) *cameras = f ? f->cameras : nullptr;
    return f ? f->n_cameras : 0u;
}

void fl_scene_file_free(fl_scene_file_t f) { delete f; }

} // extern "C"
