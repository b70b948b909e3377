// render_example.cpp -- the C++ host layer used the way filianore-core's
// caller uses Scene and PathIntegrator: build a SceneGeometry of shared
// triangles, a Scene over it, intersect rays one at a time and in batches,
// and render a film with GpuPathIntegrator.
//
//   fl_render_example [film.bin] [width height spp [standard [textures.bin]]]
//
// The scene is the procedural Cornell box (include/fl_scenegen.h), rebuilt
// here as one TriangleMesh per run of equal material so that triangle IDs
// keep the generator's order. With "standard", the walls and blocks become
// StandardSurfaceMaterials (Material::standard_surface); with a textures file
// (5 images, each uint32 width, uint32 height, then width * height RGB
// bytes) they also read ImageMapTextures at uv = ((x + z + 2) / 4,
// (y + 1) / 2). Prints one JSON line; writes the film sums
// (float32, height x width x 3) to film.bin when given. Exit status: 0 ok,
// 2 no GPU, 1 any other failure.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/fl_scenegen.h"
#include "filianore_gpu.h"

using namespace filianore::gpu;

int main(int argc, char **argv) {
    const char *film_path = argc > 1 ? argv[1] : nullptr;
    fl_gen_options opt;
    fl_gen_default_options(&opt);
    if (argc > 4) {
        opt.width = (uint32_t)atoi(argv[2]);
        opt.height = (uint32_t)atoi(argv[3]);
    }
    fl_gen_scene g;
    if (fl_gen_config(FL_GEN_CORNELL, &opt, &g) != FL_OK) {
        fprintf(stderr, "scene generator failed\n");
        return 1;
    }
    const int spp = argc > 4 ? atoi(argv[4]) : (int)g.spp;

    // SceneGeometry in the reference's shape: shared_ptr per triangle
    SceneGeometry geo;
    std::vector<size_t> mesh_of(g.n_triangles), index_in_mesh(g.n_triangles);
    for (uint32_t t = 0; t < g.n_triangles; ++t) {
        const uint32_t mat = g.tri_material ? g.tri_material[t] : 0u;
        if (geo.triangleMeshes.empty() || geo.triangleMeshes.back()->material != mat) {
            geo.triangleMeshes.push_back(std::make_shared<TriangleMesh>());
            geo.triangleMeshes.back()->material = mat;
        }
        auto tri = std::make_shared<BasicTriangle>();
        const float *v = g.vertices;
        const uint32_t *ix = g.indices + 3 * t;
        tri->v1 = Vector3f(v[3 * ix[0]], v[3 * ix[0] + 1], v[3 * ix[0] + 2]);
        tri->v2 = Vector3f(v[3 * ix[1]], v[3 * ix[1] + 1], v[3 * ix[1] + 2]);
        tri->v3 = Vector3f(v[3 * ix[2]], v[3 * ix[2] + 1], v[3 * ix[2] + 2]);
        if (argc > 6) {
            auto uv_of = [](const Vector3f &q) { return Vector2f((q.x + q.z + 2.f) / 4.f, (q.y + 1.f) / 2.f); };
            tri->uv1 = uv_of(tri->v1);
            tri->uv2 = uv_of(tri->v2);
            tri->uv3 = uv_of(tri->v3);
        }
        mesh_of[t] = geo.triangleMeshes.size() - 1;
        index_in_mesh[t] = geo.triangleMeshes.back()->triangles.size();
        geo.triangleMeshes.back()->triangles.push_back(tri);
    }
    std::vector<Material> mats;
    for (uint32_t m = 0; m < g.n_materials; ++m) {
        Material mm;
        mm.m = g.materials[m];
        mats.push_back(mm);
    }
    if (argc > 5 && strcmp(argv[5], "standard") == 0 && mats.size() >= 4) {
        // white walls: blended diffuse/specular, thin film, sheen, coat; red wall:
        // Oren-Nayar under a wavelength-dependent metal, thin film, coat; green
        // wall: Schlick-metallic blend; blocks: thin-walled translucent
        const Vector3f ks(0.9f, 0.85f, 0.8f), kt(0.8f, 0.9f, 0.95f), sheen(0.9f, 0.6f, 0.4f), coat(0.95f, 0.95f, 1.f);
        mats[0] = Material::standard_surface(0.6f, Vector3f(0.75f, 0.72f, 0.68f), 0.5f, 0.6f, ks, 0.35f, 0.3f,
                                             FL_FRESNEL_DIELECTRIC, 1.5f, 0, 0.5f, sheen, 0.6f, 0.f, kt, 0.5f, coat,
                                             1.4f, 0.5f, 410.f, 1.45f, false);
        mats[1] = Material::standard_surface(0.5f, Vector3f(0.65f, 0.12f, 0.10f), 0.3f, 0.7f, ks, 0.5f, 0.f,
                                             FL_FRESNEL_WAVELENGTH_METALLIC, 1.5f, 2, 0.f, sheen, 0.f, 0.f, kt, 0.4f,
                                             coat, 1.5f, 0.2f, 250.f, 1.33f, false);
        mats[2] = Material::standard_surface(0.5f, Vector3f(0.12f, 0.45f, 0.15f), 0.2f, 0.9f, ks, 0.55f, 0.6f,
                                             FL_FRESNEL_METALLIC, 1.5f, 4, 0.f, sheen, 0.f, 0.f, kt, 0.f, coat, 1.5f,
                                             0.f, 0.f, 1.33f, false);
        mats[3] = Material::standard_surface(0.5f, Vector3f(0.7f, 0.7f, 0.6f), 0.6f, 0.8f, ks, 0.f, 0.f,
                                             FL_FRESNEL_WAVELENGTH_METALLIC, 1.45f, 6, 0.f, sheen, 0.f, 0.9f, kt, 0.3f,
                                             coat, 1.5f, 0.4f, 0.f, 1.33f, true);
        if (argc > 6) {
            std::vector<std::shared_ptr<const ImageMapTexture>> tex;
            FILE *f = fopen(argv[6], "rb");
            for (int k = 0; f && k < 5; ++k) {
                auto t = std::make_shared<ImageMapTexture>();
                uint32_t wh[2] = {0, 0};
                if (fread(wh, 4, 2, f) != 2) break;
                t->width = wh[0];
                t->height = wh[1];
                t->rgb.resize(3ull * wh[0] * wh[1]);
                if (fread(t->rgb.data(), 1, t->rgb.size(), f) != t->rgb.size()) break;
                tex.push_back(t);
            }
            if (f) fclose(f);
            if (tex.size() != 5) {
                printf("{\"ok\": false, \"error\": \"cannot read %s\"}\n", argv[6]);
                fl_gen_free(&g);
                return 1;
            }
            mats[0].with_texture(TextureSlot::Kd, tex[0]).with_texture(TextureSlot::SheenColor, tex[1])
                .with_texture(TextureSlot::KrCoat, tex[2]);
            mats[1].with_texture(TextureSlot::Kd, tex[3]);
            mats[2].with_texture(TextureSlot::Ks, tex[4]);
            mats[3].with_texture(TextureSlot::Kd, tex[1]).with_texture(TextureSlot::Kt, tex[2]);
        }
    }
    std::vector<DiffuseAreaIlluminant> lights;
    for (uint32_t l = 0; l < g.n_lights; ++l) {
        DiffuseAreaIlluminant il;
        il.mesh = mesh_of[g.lights[l].tri];
        il.triangle = index_in_mesh[g.lights[l].tri];
        il.spd.assign(g.emission_spd + (size_t)g.lights[l].emission * FL_SPD_N,
                      g.emission_spd + (size_t)(g.lights[l].emission + 1) * FL_SPD_N);
        il.intensity = g.lights[l].intensity;
        lights.push_back(il);
    }

    Device dev(0);
    if (!dev.ok()) {
        printf("{\"ok\": false, \"error\": \"%s\"}\n", dev.error().c_str());
        fl_gen_free(&g);
        return dev.status() == FL_E_UNSUPPORTED ? 2 : 1;
    }
    Scene scene(dev, geo, mats, lights);
    if (!scene.ok()) {
        printf("{\"ok\": false, \"error\": \"%s\"}\n", scene.error().c_str());
        fl_gen_free(&g);
        return 1;
    }

    // rays through pixel centres: single-ray and batch forms must agree
    const fl_camera &gc = g.camera;
    PhysicalCamera cam(Vector3f(gc.position[0], gc.position[1], gc.position[2]),
                       Vector3f(gc.look_at[0], gc.look_at[1], gc.look_at[2]), (int)gc.width, (int)gc.height,
                       gc.vfov_deg, 0.f, 0.f, 0.f, Vector3f(gc.up[0], gc.up[1], gc.up[2]), true);
    std::vector<Ray> rays;
    for (int i = 0; i < 64; ++i) {
        Ray r;
        r.origin = Vector3f(gc.position[0], gc.position[1], gc.position[2]);
        const float a = -0.5f + (float)(i % 8) / 8.f, b = -0.5f + (float)(i / 8) / 8.f;
        r.dir = Vector3f(a * 0.7f, b * 0.7f, -1.f);
        rays.push_back(r);
    }
    std::vector<SurfaceInteraction> hits;
    std::vector<uint8_t> hit, occ;
    int mismatches = 0, n_hit = 0, n_occ = 0;
    if (!scene.intersect(rays, hits, hit) || !scene.intersect_p(rays, occ)) {
        printf("{\"ok\": false, \"error\": \"%s\"}\n", scene.error().c_str());
        fl_gen_free(&g);
        return 1;
    }
    for (size_t i = 0; i < rays.size(); ++i) {
        SurfaceInteraction s;
        const bool h1 = scene.intersect(rays[i], &s);
        const bool o1 = scene.intersect_p(rays[i]);
        n_hit += hit[i];
        n_occ += occ[i];
        if (h1 != (hit[i] != 0) || o1 != (occ[i] != 0) || (h1 && (s.primID != hits[i].primID || s.t != hits[i].t)))
            ++mismatches;
        if ((hit[i] != 0) != (occ[i] != 0)) ++mismatches;
    }

    RenderParams rp;
    rp.samples = spp;
    rp.diffuseRayDepth = rp.reflectionRayDepth = rp.transmissionRayDepth = g.depth;
    GpuPathIntegrator integrator(rp);
    integrator.prepare_the_renderer(scene);
    Film film;
    const auto t0 = std::chrono::steady_clock::now();
    const bool ok = integrator.render(scene, cam, film, 7);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!ok) {
        printf("{\"ok\": false, \"error\": \"%s\"}\n", integrator.error().c_str());
        fl_gen_free(&g);
        return 1;
    }
    double mean[3] = {0, 0, 0};
    for (size_t i = 0; i < film.xyz_sum.size(); ++i) mean[i % 3] += film.xyz_sum[i];
    const double npx = (double)film.width * film.height * (double)film.samples_per_pixel;
    if (film_path) {
        FILE *f = fopen(film_path, "wb");
        if (!f || fwrite(film.xyz_sum.data(), 4, film.xyz_sum.size(), f) != film.xyz_sum.size()) {
            fprintf(stderr, "cannot write %s\n", film_path);
            if (f) fclose(f);
            fl_gen_free(&g);
            return 1;
        }
        fclose(f);
    }
    printf("{\"ok\": true, \"triangles\": %u, \"meshes\": %zu, \"width\": %d, \"height\": %d, \"spp\": %d, "
           "\"ray_hits\": %d, \"ray_occluded\": %d, \"mismatches\": %d, \"mean_xyz\": [%.6g, %.6g, %.6g], "
           "\"seconds\": %.4f, \"msamples_per_s\": %.2f}\n",
           g.n_triangles, geo.triangleMeshes.size(), film.width, film.height, spp, n_hit, n_occ, mismatches,
           mean[0] / npx, mean[1] / npx, mean[2] / npx, secs, npx / secs / 1e6);
    fl_gen_free(&g);
    return mismatches == 0 ? 0 : 1;
}
