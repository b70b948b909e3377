// filianore_gpu.h -- C++ host layer of the MI355X render path, over the C ABI
// of libfl_gpu.so (include/fl_gpu.h). Header-only; link with -lfl_gpu.
//
// It mirrors the reference's render-path interface, with the same names and
// argument meanings, so a caller of filianore-core finds what it uses:
//
//   SceneGeometry / TriangleMesh / BasicTriangle   bvh/bvh.h:9-19
//   Scene(geometry, illuminants), intersect,        core/scene.h:12-31,
//     intersect_p                                    core/scene.cpp:7-31
//   Ray {origin, dir, tMin, tMax}                    maths/geometry.h (Ray)
//   RenderParams {samples, diffuseRayDepth,          core/renderparams.h:5-10
//     reflectionRayDepth, transmissionRayDepth}
//   PhysicalCamera(pos, lookAt, xRes, yRes, vFOV,    cameras/physical.h:11-43
//     aperture, t0, t1, up, flipCameraDirection)
//   PathIntegrator(renderParams, rrThreshold)        integrators/pathintegrator.h:9-22
//     -> GpuPathIntegrator: prepare_the_renderer + a batch render() in place
//        of the per-ray Li(), which the GPU runs for every camera sample
//
// Differences, each deliberate:
//   - errors: like the reference, nothing throws; calls return bool (hit /
//     success) and error() holds the message of the last failure.
//   - Scene is move-only (the reference's implicitly copyable BVH releases
//     its Embree scene twice when a Scene is copied).
//   - a mesh carries the index of its material, and an illuminant names the
//     triangle it covers: the reference's Embree path loses both (its hits
//     carry no primitive), so the render path needs them given explicitly.
//   - PhysicalCamera is a pinhole (aperture 0) with flipCameraDirection =
//     true, the orientation that looks at lookAt; other values are rejected.
//   - single-ray intersect()/intersect_p() are exact but launch one kernel
//     per call; batch callers use the vector forms.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/fl_gpu.h"

namespace filianore {
namespace gpu {

struct Vector3f {
    float x = 0.f, y = 0.f, z = 0.f;
    Vector3f() = default;
    Vector3f(float a, float b, float c) : x(a), y(b), z(c) {}
};

struct Vector2f {
    float x = 0.f, y = 0.f;
    Vector2f() = default;
    Vector2f(float a, float b) : x(a), y(b) {}
};

// ---- scene input (bvh/bvh.h:9-19) ----------------------------------------
// uv1..uv3: the vertices' texture coordinates (the Vertex::uv of
// shapes/triangle.h); a scene where every triangle leaves them at (0, 0)
// has no uvs, and every hit reads a texture at uv (0, 0) as the reference's
// Embree path does.
struct BasicTriangle {
    Vector3f v1, v2, v3;
    Vector2f uv1, uv2, uv3;
};

struct TriangleMesh {
    std::vector<std::shared_ptr<BasicTriangle>> triangles;
    uint32_t material = 0;   // index into the material list given to Scene
};

struct SceneGeometry {
    std::vector<std::shared_ptr<TriangleMesh>> triangleMeshes;
};

// ---- materials the render path implements --------------------------------
// ImageMapTexture (shading/textures/imagemap.h:8-23): the reference loads the
// file with stb_image; here the caller passes the decoded 8-bit RGB image,
// rows top to bottom. An empty image is black, as one that failed to load.
struct ImageMapTexture {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> rgb;   // width * height * 3
};

// which spectrum of a material an image texture replaces
enum class TextureSlot { Kd = 0, Ks = 1, SheenColor = 2, Kt = 3, KrCoat = 4 };

struct Material {
    fl_material m{};
    // image textures in place of constant colours (Lambert: Kd only)
    std::shared_ptr<const ImageMapTexture> textures[5];
    Material &with_texture(TextureSlot which, std::shared_ptr<const ImageMapTexture> t) {
        textures[(int)which] = std::move(t);
        return *this;
    }
    // LambertMaterial(kd), shading/materials/lambert.cpp:9-17
    static Material lambert(Vector3f kd) {
        Material r;
        r.m.kind = FL_MAT_LAMBERT;
        r.m.rgb[0] = kd.x; r.m.rgb[1] = kd.y; r.m.rgb[2] = kd.z;
        r.m.weight = 1.f;
        return r;
    }
    // StandardSurface metal branch: ksweight, forge metal, ksroughness,
    // ksanisotropic (shading/materials/standardsurface.cpp:101-125)
    static Material conductor(int metal, float roughness, float weight = 1.f, float anisotropic = 0.f) {
        Material r;
        r.m.kind = FL_MAT_CONDUCTOR;
        r.m.metal = metal;
        r.m.roughness = roughness;
        r.m.weight = weight;
        r.m.anisotropic = anisotropic;
        return r;
    }
    // StandardSurface smooth dielectric (standardsurface.cpp:77-81,113-117);
    // cauchyB != 0 makes eta wavelength dependent (A = ior)
    static Material dielectric(float ior, Vector3f ks = {1.f, 1.f, 1.f}, Vector3f kt = {1.f, 1.f, 1.f},
                               float cauchyB = 0.f) {
        Material r;
        r.m.kind = FL_MAT_DIELECTRIC;
        r.m.ior = ior;
        r.m.cauchy_b = cauchyB;
        r.m.rgb[0] = ks.x; r.m.rgb[1] = ks.y; r.m.rgb[2] = ks.z;
        r.m.rgb_t[0] = kt.x; r.m.rgb_t[1] = kt.y; r.m.rgb_t[2] = kt.z;
        r.m.weight = 1.f;
        return r;
    }
    // StandardSurfaceMaterial (shading/materials/standardsurface.h:14-37), its
    // constructor's parameters in their order, every texture constant (colours
    // as reflectance RGB); the bump map is not supported. fresnelType is
    // FL_FRESNEL_* (FresnelType), metalType a forge metal.
    static Material standard_surface(float kdweight, Vector3f kd, float kdroughness, float ksweight, Vector3f ks,
                                     float ksroughness, float ksanisotropic, int fresnelType, float ksIOR,
                                     int metalType, float sheenweight, Vector3f sheenColor, float sheenroughness,
                                     float ktweight, Vector3f kt, float krcoatweight, Vector3f krcoat,
                                     float krcoatior, float krcoatgloss, float thinFilmThickness,
                                     float thinFilmIOR, bool thinWalled) {
        Material r;
        r.m.kind = FL_MAT_STANDARD;
        r.has_standard = true;
        fl_standard_surface &p = r.standard;
        auto rgb = [](float *o, const Vector3f &v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; };
        p.kd_weight = kdweight; rgb(p.kd, kd); p.kd_roughness = kdroughness;
        p.ks_weight = ksweight; rgb(p.ks, ks); p.ks_roughness = ksroughness; p.ks_anisotropic = ksanisotropic;
        p.fresnel_type = fresnelType; p.ks_ior = ksIOR; p.metal = metalType;
        p.sheen_weight = sheenweight; rgb(p.sheen_color, sheenColor); p.sheen_roughness = sheenroughness;
        p.kt_weight = ktweight; rgb(p.kt, kt);
        p.coat_weight = krcoatweight; rgb(p.coat, krcoat); p.coat_ior = krcoatior; p.coat_gloss = krcoatgloss;
        p.thin_film_thickness = thinFilmThickness; p.thin_film_ior = thinFilmIOR;
        p.thin_walled = thinWalled ? 1 : 0;
        return r;
    }
    // a primitive without material: paths pass through it
    // (integrators/pathintegrator.cpp:33-38)
    static Material none() {
        Material r;
        r.m.kind = FL_MAT_NONE;
        return r;
    }
    bool has_standard = false;
    fl_standard_surface standard{};   // FL_MAT_STANDARD: the material's parameters
};

// DiffuseAreaIlluminant over one triangle (illuminants/diffusearea.cpp:7-39).
// spd: FL_SPD_N samples on 360..830 nm (1 nm), scaled by intensity.
struct DiffuseAreaIlluminant {
    size_t mesh = 0, triangle = 0;   // the emitting triangle of the geometry
    std::vector<float> spd;
    float intensity = 1.f;
    std::vector<float> shadowColor;  // Illuminant::shadowColor (FL_SPD_N samples); empty = black
};

// Delta illuminants. The reference builds them from a lightToWorld Transform;
// here the caller passes what the constructors derive from it: the light's
// world position (lightToWorld applied to the origin) and its world
// direction (vector_transform of dirIllum).
// PointIlluminant (illuminants/point.cpp:6-10)
struct PointIlluminant {
    Vector3f position;
    std::vector<float> color;        // FL_SPD_N samples on 360..830 nm
    float intensity = 1.f;
    std::vector<float> shadowColor;  // empty = black
};

// SpotIlluminant (illuminants/spot.cpp:6-21): coneAngle / penumbraAngle in
// degrees unless angleInRadians, halved when usehalfAngles, as the reference
// constructor does; light is 0 outside the cone, 1 inside the penumbra
// angle, smoothstep between.
struct SpotIlluminant {
    Vector3f position, direction;
    float coneAngle = 45.f, penumbraAngle = 40.f;
    bool angleInRadians = false, usehalfAngles = false;
    std::vector<float> color;
    float intensity = 1.f;
    std::vector<float> shadowColor;
    float cos_cone() const { return std::cos(angle(coneAngle)); }
    float cos_penumbra() const { return std::cos(angle(penumbraAngle)); }

private:
    float angle(float a) const {
        float r = angleInRadians ? a : a * (3.14159265358979323846f / 180.f);
        return usehalfAngles ? r / 2.f : r;
    }
};

// DirectionalIlluminant (illuminants/directional.cpp:6-13): direction is the
// way the light travels; the light sits at -direction, beyond the world.
struct DirectionalIlluminant {
    Vector3f direction;
    std::vector<float> color;
    float intensity = 1.f;
    std::vector<float> shadowColor;
};

// Scene illuminants besides the area lights, sampled after them.
struct DeltaIlluminants {
    std::vector<PointIlluminant> points;
    std::vector<SpotIlluminant> spots;
    std::vector<DirectionalIlluminant> directionals;
};

// Planck spectrum at temperature T normalized to a peak of 1 over 360..830 nm.
inline std::vector<float> blackbody_spd(float kelvin) {
    std::vector<float> s(FL_SPD_N);
    const double c = 299792458.0, h = 6.62606957e-34, kb = 1.3806488e-23;
    double peak = 0.0;
    for (int i = 0; i < FL_SPD_N; ++i) {
        double l = (360.0 + i) * 1e-9;
        double v = (2.0 * h * c * c) / (std::pow(l, 5) * (std::exp((h * c) / (l * kb * kelvin)) - 1.0));
        s[i] = (float)v;
        peak = std::max(peak, v);
    }
    for (auto &v : s) v = (float)(v / peak);
    return s;
}

// ---- rays and hits ---------------------------------------------------------
struct Ray {
    Vector3f origin, dir;
    float tMin = 0.f;
    float tMax = std::numeric_limits<float>::infinity();
};

// The hit record of BVH::intersect (bvh/bvh.cpp:49-84): t, point, geometric
// normal normalize((v2-v1) x (v3-v1)), Embree barycentrics, triangle ID.
struct SurfaceInteraction {
    float t = 0.f;
    Vector3f p, n;
    float u = 0.f, v = 0.f;
    uint32_t primID = FL_NO_HIT;
};

// ---- device ----------------------------------------------------------------
class Device {
public:
    explicit Device(int device = 0) { status_ = fl_ctx_create(device, &ctx_); }
    ~Device() {
        if (ctx_) fl_ctx_destroy(ctx_);
    }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    bool ok() const { return ctx_ != nullptr; }
    int status() const { return status_; }
    std::string error() const {
        if (!ctx_) return "no usable GPU (fl_ctx_create returned " + std::to_string(status_) + ")";
        return fl_last_error(ctx_);
    }
    fl_ctx_t handle() const { return ctx_; }

private:
    fl_ctx_t ctx_ = nullptr;
    int status_ = FL_OK;
};

// ---- Scene (core/scene.h:12-31) --------------------------------------------
class Scene {
public:
    Scene(Device &dev, const SceneGeometry &geometry, const std::vector<Material> &materials,
          const std::vector<DiffuseAreaIlluminant> &illuminants, const DeltaIlluminants &deltas = DeltaIlluminants())
        : ctx_(dev.handle()) {
        if (!ctx_) {
            err_ = dev.error();
            return;
        }
        // flattened (mesh, triangle) order = triangle ID (bvh/bvh.cpp:12-16,41)
        std::vector<size_t> first(geometry.triangleMeshes.size() + 1, 0);
        for (size_t m = 0; m < geometry.triangleMeshes.size(); ++m)
            first[m + 1] = first[m] + geometry.triangleMeshes[m]->triangles.size();
        const size_t n = first.back();
        std::vector<float> verts(n * 9), uvs(n * 6);
        std::vector<uint32_t> idx(n * 3), tmat(n);
        bool has_uv = false;
        for (size_t m = 0; m < geometry.triangleMeshes.size(); ++m) {
            const TriangleMesh &mesh = *geometry.triangleMeshes[m];
            for (size_t t = 0; t < mesh.triangles.size(); ++t) {
                const size_t id = first[m] + t;
                const BasicTriangle &tr = *mesh.triangles[t];
                const Vector3f v[3] = {tr.v1, tr.v2, tr.v3};
                const Vector2f uv[3] = {tr.uv1, tr.uv2, tr.uv3};
                for (int k = 0; k < 3; ++k) {
                    verts[id * 9 + 3 * k + 0] = v[k].x;
                    verts[id * 9 + 3 * k + 1] = v[k].y;
                    verts[id * 9 + 3 * k + 2] = v[k].z;
                    uvs[id * 6 + 2 * k + 0] = uv[k].x;
                    uvs[id * 6 + 2 * k + 1] = uv[k].y;
                    has_uv = has_uv || uv[k].x != 0.f || uv[k].y != 0.f;
                    idx[id * 3 + k] = (uint32_t)(id * 3 + k);
                }
                tmat[id] = mesh.material;
            }
        }
        std::vector<fl_material> mats;
        std::vector<fl_standard_surface> standard;
        std::vector<fl_texture> textures;
        std::vector<const ImageMapTexture *> tex_seen;
        // 1 + index of a texture in `textures` (each image once)
        auto tex_ref = [&](const std::shared_ptr<const ImageMapTexture> &t) -> uint32_t {
            if (!t) return 0u;
            for (size_t j = 0; j < tex_seen.size(); ++j)
                if (tex_seen[j] == t.get()) return (uint32_t)j + 1u;
            fl_texture ft{};
            if (t->rgb.size() >= 3ull * t->width * t->height) {
                ft.width = t->width;
                ft.height = t->height;
                ft.rgb = t->rgb.data();
            }
            tex_seen.push_back(t.get());
            textures.push_back(ft);
            return (uint32_t)textures.size();
        };
        for (const Material &m : materials) {
            mats.push_back(m.m);
            if (m.m.kind == FL_MAT_LAMBERT) mats.back().texture = tex_ref(m.textures[(int)TextureSlot::Kd]);
            if (m.has_standard) {
                mats.back().standard = (uint32_t)standard.size();
                standard.push_back(m.standard);
                fl_standard_surface &ss = standard.back();
                ss.kd_texture = tex_ref(m.textures[(int)TextureSlot::Kd]);
                ss.ks_texture = tex_ref(m.textures[(int)TextureSlot::Ks]);
                ss.sheen_texture = tex_ref(m.textures[(int)TextureSlot::SheenColor]);
                ss.kt_texture = tex_ref(m.textures[(int)TextureSlot::Kt]);
                ss.coat_texture = tex_ref(m.textures[(int)TextureSlot::KrCoat]);
            }
        }
        std::vector<fl_area_light> lights;
        std::vector<fl_delta_light> delta;
        std::vector<float> spd;
        // appends an SPD to the emission table; returns its index, or -1 if
        // it is not FL_SPD_N samples
        auto add_spd = [&](const std::vector<float> &v) -> long {
            if (v.size() != FL_SPD_N) return -1;
            spd.insert(spd.end(), v.begin(), v.end());
            return (long)(spd.size() / FL_SPD_N) - 1;
        };
        auto shadow_of = [&](const std::vector<float> &v, bool &bad) -> uint32_t {
            if (v.empty()) return 0u;
            const long e = add_spd(v);
            bad = bad || e < 0;
            return (uint32_t)(e + 1);
        };
        for (size_t i = 0; i < illuminants.size(); ++i) {
            const DiffuseAreaIlluminant &il = illuminants[i];
            bool bad = il.mesh >= geometry.triangleMeshes.size() ||
                       il.triangle >= geometry.triangleMeshes[il.mesh]->triangles.size();
            fl_area_light l{};
            const long e = bad ? -1 : add_spd(il.spd);
            l.shadow_emission = shadow_of(il.shadowColor, bad);
            if (bad || e < 0) {
                err_ = "illuminant " + std::to_string(i) + ": no such triangle, or an spd is not 471 samples";
                return;
            }
            l.tri = (uint32_t)(first[il.mesh] + il.triangle);
            l.emission = (uint32_t)e;
            l.intensity = il.intensity;
            lights.push_back(l);
        }
        auto add_delta = [&](int type, const Vector3f &pos, const Vector3f &dir, const std::vector<float> &color,
                             float intensity, float cc, float cp, const std::vector<float> &shadow) {
            bool bad = false;
            fl_delta_light l{};
            l.type = type;
            l.position[0] = pos.x; l.position[1] = pos.y; l.position[2] = pos.z;
            l.direction[0] = dir.x; l.direction[1] = dir.y; l.direction[2] = dir.z;
            const long e = add_spd(color);
            l.shadow_emission = shadow_of(shadow, bad);
            if (bad || e < 0) return false;
            l.emission = (uint32_t)e;
            l.intensity = intensity;
            l.cos_cone = cc;
            l.cos_penumbra = cp;
            delta.push_back(l);
            return true;
        };
        bool ok = true;
        for (const PointIlluminant &l : deltas.points)
            ok = ok && add_delta(FL_LIGHT_POINT, l.position, Vector3f(), l.color, l.intensity, 0.f, 0.f, l.shadowColor);
        for (const SpotIlluminant &l : deltas.spots)
            ok = ok && add_delta(FL_LIGHT_SPOT, l.position, l.direction, l.color, l.intensity, l.cos_cone(),
                                 l.cos_penumbra(), l.shadowColor);
        for (const DirectionalIlluminant &l : deltas.directionals)
            ok = ok && add_delta(FL_LIGHT_DIRECTIONAL, Vector3f(), l.direction, l.color, l.intensity, 0.f, 0.f,
                                 l.shadowColor);
        if (!ok) {
            err_ = "delta illuminant: a color or shadowColor is not 471 samples";
            return;
        }
        fl_scene_desc d{};
        d.vertices = verts.data();
        d.n_vertices = (uint32_t)(n * 3);
        d.indices = idx.data();
        d.n_triangles = (uint32_t)n;
        d.tri_material = tmat.data();
        d.materials = mats.data();
        d.n_materials = (uint32_t)mats.size();
        d.lights = lights.data();
        d.n_lights = (uint32_t)lights.size();
        d.emission_spd = spd.empty() ? nullptr : spd.data();
        d.n_emission = (uint32_t)(spd.size() / FL_SPD_N);
        d.delta_lights = delta.empty() ? nullptr : delta.data();
        d.n_delta_lights = (uint32_t)delta.size();
        d.standard_surfaces = standard.empty() ? nullptr : standard.data();
        d.n_standard_surfaces = (uint32_t)standard.size();
        d.textures = textures.empty() ? nullptr : textures.data();
        d.n_textures = (uint32_t)textures.size();
        d.uvs = has_uv ? uvs.data() : nullptr;
        if (fl_scene_create(ctx_, &d, &scene_) != FL_OK) {
            err_ = fl_last_error(ctx_);
            scene_ = nullptr;
        }
    }
    ~Scene() {
        if (scene_) fl_scene_destroy(scene_);
    }
    Scene(Scene &&o) noexcept : ctx_(o.ctx_), scene_(o.scene_), err_(std::move(o.err_)) { o.scene_ = nullptr; }
    Scene &operator=(Scene &&o) noexcept {
        if (this != &o) {
            if (scene_) fl_scene_destroy(scene_);
            ctx_ = o.ctx_;
            scene_ = o.scene_;
            err_ = std::move(o.err_);
            o.scene_ = nullptr;
        }
        return *this;
    }
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;

    bool ok() const { return scene_ != nullptr; }
    const std::string &error() const { return err_; }
    fl_scene_t handle() const { return scene_; }

    // Scene::intersect (core/scene.cpp:25-27): closest hit in (tMin, tMax]
    bool intersect(const Ray &ray, SurfaceInteraction *isect) const {
        std::vector<SurfaceInteraction> out;
        std::vector<uint8_t> hit;
        if (!intersect(std::vector<Ray>{ray}, out, hit)) return false;
        if (hit[0] && isect) *isect = out[0];
        return hit[0] != 0;
    }
    // Scene::intersect_p (core/scene.cpp:29-31): any hit
    bool intersect_p(const Ray &ray) const {
        std::vector<uint8_t> occ;
        return intersect_p(std::vector<Ray>{ray}, occ) && occ[0] != 0;
    }

    // batch forms; return false (with error()) on a failed call
    bool intersect(const std::vector<Ray> &rays, std::vector<SurfaceInteraction> &out,
                   std::vector<uint8_t> &hit) const {
        std::vector<fl_hit> h;
        if (!trace(rays, FL_CLOSEST_HIT, h)) return false;
        out.assign(rays.size(), SurfaceInteraction());
        hit.assign(rays.size(), 0);
        for (size_t i = 0; i < rays.size(); ++i) {
            if (h[i].tri == FL_NO_HIT) continue;
            hit[i] = 1;
            SurfaceInteraction &s = out[i];
            const Ray &r = rays[i];
            s.t = h[i].t;
            s.u = h[i].u;
            s.v = h[i].v;
            s.primID = h[i].tri;
            s.p = Vector3f(r.origin.x + r.dir.x * s.t, r.origin.y + r.dir.y * s.t, r.origin.z + r.dir.z * s.t);
        }
        return true;
    }
    bool intersect_p(const std::vector<Ray> &rays, std::vector<uint8_t> &occluded) const {
        std::vector<fl_hit> h;
        if (!trace(rays, FL_ANY_HIT, h)) return false;
        occluded.assign(rays.size(), 0);
        for (size_t i = 0; i < rays.size(); ++i) occluded[i] = h[i].tri != FL_NO_HIT;
        return true;
    }

private:
    bool trace(const std::vector<Ray> &rays, uint32_t flags, std::vector<fl_hit> &h) const {
        if (!scene_) return false;
        std::vector<fl_ray> r(rays.size());
        for (size_t i = 0; i < rays.size(); ++i)
            r[i] = fl_ray{rays[i].origin.x, rays[i].origin.y, rays[i].origin.z, rays[i].tMin,
                          rays[i].dir.x,    rays[i].dir.y,    rays[i].dir.z,    rays[i].tMax};
        h.resize(rays.size());
        if (fl_intersect(scene_, r.data(), h.data(), (uint32_t)r.size(), flags, nullptr) != FL_OK) {
            const_cast<std::string &>(err_) = fl_last_error(ctx_);
            return false;
        }
        return true;
    }

    fl_ctx_t ctx_ = nullptr;
    fl_scene_t scene_ = nullptr;
    std::string err_;
};

// ---- RenderParams (core/renderparams.h:5-10) --------------------------------
struct RenderParams {
    int samples = 16;
    int diffuseRayDepth = 8;
    int reflectionRayDepth = 8;
    int transmissionRayDepth = 8;
};

// ---- PhysicalCamera (cameras/physical.h:11-43) -------------------------------
class PhysicalCamera {
public:
    PhysicalCamera(const Vector3f &camPosition, const Vector3f &camLookAt, int xRes, int yRes, float vFOV,
                   float aperture, float t0, float t1, const Vector3f &camUp = Vector3f(0.f, 1.f, 0.f),
                   bool flipCameraDirection = true) {
        cam_.position[0] = camPosition.x; cam_.position[1] = camPosition.y; cam_.position[2] = camPosition.z;
        cam_.look_at[0] = camLookAt.x; cam_.look_at[1] = camLookAt.y; cam_.look_at[2] = camLookAt.z;
        cam_.up[0] = camUp.x; cam_.up[1] = camUp.y; cam_.up[2] = camUp.z;
        cam_.vfov_deg = vFOV;
        cam_.width = xRes > 0 ? (uint32_t)xRes : 0u;
        cam_.height = yRes > 0 ? (uint32_t)yRes : 0u;
        valid_ = aperture == 0.f && flipCameraDirection && xRes > 0 && yRes > 0 && t0 == 0.f && t1 == 0.f;
    }
    bool valid() const { return valid_; }
    const fl_camera &desc() const { return cam_; }

private:
    fl_camera cam_{};
    bool valid_ = false;
};

// Per-pixel sums of CIE XYZ over all samples (and of their squares).
struct Film {
    int width = 0, height = 0;
    uint64_t samples_per_pixel = 0;
    std::vector<float> xyz_sum, xyz_sqsum;   // height * width * 3, row 0 = top
    Vector3f mean_xyz(int x, int y) const {
        const size_t i = ((size_t)y * width + x) * 3;
        const float n = (float)samples_per_pixel;
        return Vector3f(xyz_sum[i] / n, xyz_sum[i + 1] / n, xyz_sum[i + 2] / n);
    }
};

// ---- GpuPathIntegrator (integrators/pathintegrator.h:9-22) -------------------
class GpuPathIntegrator {
public:
    explicit GpuPathIntegrator(const RenderParams &renderParams, float rrThreshold = 1.f)
        : renderParams(renderParams), rrThreshold_(rrThreshold) {}

    // nothing to precompute: tables and the BVH are built with the Scene
    void prepare_the_renderer(const Scene &) {}

    // Li for every camera sample of the film: spp = renderParams.samples.
    // parity = true evaluates the reference's 60-bin spectra on 400..700 nm.
    bool render(const Scene &scene, const PhysicalCamera &camera, Film &film, uint64_t seed = 1,
                bool parity = false) {
        if (!scene.ok()) {
            err_ = "scene not built: " + scene.error();
            return false;
        }
        if (!camera.valid()) {
            err_ = "camera: only a pinhole (aperture 0, t0 = t1 = 0) with flipCameraDirection = true is supported";
            return false;
        }
        if (renderParams.samples <= 0) {
            err_ = "renderParams.samples must be positive";
            return false;
        }
        const fl_camera &c = camera.desc();
        film.width = (int)c.width;
        film.height = (int)c.height;
        film.samples_per_pixel = (uint64_t)renderParams.samples;
        film.xyz_sum.assign((size_t)c.width * c.height * 3, 0.f);
        film.xyz_sqsum.assign((size_t)c.width * c.height * 3, 0.f);
        fl_render_params p{};
        p.spp = (uint32_t)renderParams.samples;
        p.diffuse_depth = renderParams.diffuseRayDepth;
        p.reflection_depth = renderParams.reflectionRayDepth;
        p.transmission_depth = renderParams.transmissionRayDepth;
        p.rr_threshold = rrThreshold_;
        p.rr_start = 3;   // RR after bounce 3 (pathintegrator.cpp:77-83)
        p.spectral_mode = parity ? FL_SPECTRAL_PARITY : FL_SPECTRAL_HERO;
        p.flags = FL_RENDER_GRAPH | (parity ? 0u : (uint32_t)FL_RENDER_VNDF_PDF) | extra_flags_;
        p.seed = seed;
        fl_film f{};
        f.xyz_sum = film.xyz_sum.data();
        f.xyz_sqsum = film.xyz_sqsum.data();
        f.width = c.width;
        f.height = c.height;
        if (fl_render(scene.handle(), &c, &p, &f, &stats_) != FL_OK) {
            err_ = "fl_render failed";
            return false;
        }
        return true;
    }

    const fl_stats &stats() const { return stats_; }
    const std::string &error() const { return err_; }

    RenderParams renderParams;

protected:
    uint32_t extra_flags_ = 0;   // FL_RENDER_* added to every render

private:
    float rrThreshold_;
    fl_stats stats_{};
    std::string err_;
};

// ---- GpuNormalIntegrator (integrators/normalintegrator.h) --------------------
// Li = from_reflectance_rgb(isect.n) of the first hit, black on a miss
// (normalintegrator.cpp:11-26); renderParams.samples samples per pixel.
class GpuNormalIntegrator : public GpuPathIntegrator {
public:
    explicit GpuNormalIntegrator(const RenderParams &renderParams) : GpuPathIntegrator(renderParams) {
        extra_flags_ = FL_RENDER_NORMALS;
    }
};

} // namespace gpu
} // namespace filianore
