// rtc_example.cpp -- code written against Embree 4 in the call pattern of
// filianore-core's BVH (bvh/bvh.cpp:7-88: one RTCGeometry per triangle with
// a 3-vertex buffer and index (0,1,2), rtcCommitScene, rtcIntersect1 per ray,
// geomID as the primitive ID, Ng normalized for the shading normal), compiled
// unchanged against include/fl_rtc.h and run on the GPU LBVH.
//
//   fl_rtc_example [n_rays]
//
// Checks, on the displaced-icosphere scene at subdivision 2 (2,562 triangles):
//   - per-triangle geometries: every rtcIntersect1 equals fl_intersect on the
//     same triangles given as one indexed buffer (t, u, v bit for bit;
//     geomID = triangle, primID = 0) and Ng equals (v1-v0) x (v2-v0);
//   - rtcOccluded1 sets tfar = -inf exactly for the rays that hit;
//   - meshes of varying size with a 16-byte vertex stride: (geomID, primID)
//     address the same triangle, t, u, v bit for bit;
//   - a triangle given with the opposite winding: Ng flips;
//   - errors: a query before rtcCommitScene and a wrong buffer format report
//     RTC_ERROR_INVALID_OPERATION / RTC_ERROR_INVALID_ARGUMENT.
// Prints one JSON line. Exit status: 0 ok, 2 no GPU, 1 any other failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/fl_gpu.h"
#include "../../include/fl_rtc.h"
#include "../../include/fl_scenegen.h"

namespace {

struct Vec3 {
    float x, y, z;
};

struct Tri {
    Vec3 v1, v2, v3;
};

// The reference's BVH class shape: add triangles, finalise, intersect.
class EmbreeBvh {
  public:
    explicit EmbreeBvh(RTCDevice d) : device(d), scene(rtcNewScene(d)) {}
    ~EmbreeBvh() { rtcReleaseScene(scene); }

    void add_triangle(const Tri &t) {
        RTCGeometry geometry = rtcNewGeometry(device, RTC_GEOMETRY_TYPE_TRIANGLE);
        Vec3 *verts = (Vec3 *)rtcSetNewGeometryBuffer(geometry, RTC_BUFFER_TYPE_VERTEX, 0, RTC_FORMAT_FLOAT3,
                                                      sizeof(Vec3), 3);
        verts[0] = t.v1;
        verts[1] = t.v2;
        verts[2] = t.v3;
        unsigned *index = (unsigned *)rtcSetNewGeometryBuffer(geometry, RTC_BUFFER_TYPE_INDEX, 0, RTC_FORMAT_UINT3,
                                                             3 * sizeof(unsigned), 1);
        index[0] = 0;
        index[1] = 1;
        index[2] = 2;
        rtcCommitGeometry(geometry);
        rtcAttachGeometry(scene, geometry);
        rtcReleaseGeometry(geometry);
    }

    // a mesh of several triangles with its own vertex buffer (16-byte stride:
    // strided buffers must be read as such); reverse = opposite winding
    void add_mesh(const float *vertices, const uint32_t *indices, uint32_t first_tri, uint32_t n_tris,
                  bool reverse = false) {
        RTCGeometry geometry = rtcNewGeometry(device, RTC_GEOMETRY_TYPE_TRIANGLE);
        float *verts = (float *)rtcSetNewGeometryBuffer(geometry, RTC_BUFFER_TYPE_VERTEX, 0, RTC_FORMAT_FLOAT3,
                                                        16, 3 * n_tris);
        unsigned *index = (unsigned *)rtcSetNewGeometryBuffer(geometry, RTC_BUFFER_TYPE_INDEX, 0, RTC_FORMAT_UINT3,
                                                             12, n_tris);
        for (uint32_t t = 0; t < n_tris; ++t)
            for (int k = 0; k < 3; ++k) {
                const uint32_t src = indices[3 * (first_tri + t) + k];
                std::memcpy(verts + 4 * (3 * t + k), vertices + 3 * src, 12);
                index[3 * t + k] = 3 * t + (reverse ? 2 - k : k);
            }
        rtcCommitGeometry(geometry);
        rtcAttachGeometry(scene, geometry);
        rtcReleaseGeometry(geometry);
    }

    void finalise() { rtcCommitScene(scene); }

    bool intersect(const fl_ray &ray, RTCRayHit *out) const {
        RTCRayHit rayhit;
        rayhit.ray.org_x = ray.ox;
        rayhit.ray.org_y = ray.oy;
        rayhit.ray.org_z = ray.oz;
        rayhit.ray.dir_x = ray.dx;
        rayhit.ray.dir_y = ray.dy;
        rayhit.ray.dir_z = ray.dz;
        rayhit.ray.tnear = ray.tmin;
        rayhit.ray.tfar = ray.tmax;
        rayhit.hit.geomID = RTC_INVALID_GEOMETRY_ID;
        rayhit.hit.primID = RTC_INVALID_GEOMETRY_ID;

        RTCRayQueryContext context;
        rtcInitRayQueryContext(&context);

        rtcIntersect1(scene, &rayhit);
        *out = rayhit;
        return rayhit.hit.geomID != RTC_INVALID_GEOMETRY_ID && rayhit.hit.primID != RTC_INVALID_GEOMETRY_ID;
    }

    bool occluded(const fl_ray &ray) const {
        RTCRay r;
        r.org_x = ray.ox;
        r.org_y = ray.oy;
        r.org_z = ray.oz;
        r.dir_x = ray.dx;
        r.dir_y = ray.dy;
        r.dir_z = ray.dz;
        r.tnear = ray.tmin;
        r.tfar = ray.tmax;
        rtcOccluded1(scene, &r);
        return r.tfar == -INFINITY;
    }

    RTCDevice device;
    RTCScene scene;
};

uint64_t rng_state = 0x243F6A8885A308D3ull;
float uniform01() {
    rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
    return (float)((rng_state >> 40) * (1.0 / 16777216.0));
}

Vec3 cross_edges(const float *a, const float *b, const float *c) {
    const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    return {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
}

// Ng within float rounding of the expected normal n (scaled by s = +-1)
bool ng_close(const Vec3 &n, float s, const RTCHit &h) {
    const float len = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    const float tol = 1e-5f * len;
    return std::fabs(s * n.x - h.Ng_x) <= tol && std::fabs(s * n.y - h.Ng_y) <= tol && std::fabs(s * n.z - h.Ng_z) <= tol;
}

int fail(const char *msg) {
    printf("{\"ok\": false, \"error\": \"%s\"}\n", msg);
    return 1;
}

} // namespace

int main(int argc, char **argv) {
    const uint32_t n_rays = argc > 1 ? (uint32_t)atoi(argv[1]) : 2048u;
    RTCDevice device = rtcNewDevice(nullptr);
    if (!device) {
        printf("{\"ok\": false, \"error\": \"no usable GPU\"}\n");
        return 2;
    }

    fl_gen_options opt;
    fl_gen_default_options(&opt);
    opt.subdiv = 2;
    fl_gen_scene g;
    if (fl_gen_config(FL_GEN_SPHERES, &opt, &g) != FL_OK) return fail("scene generator failed");
    const uint32_t nt = g.n_triangles;

    // the same triangles as one indexed buffer, straight through the C ABI
    fl_ctx_t ctx = nullptr;
    if (fl_ctx_create(0, &ctx) != FL_OK) return fail("fl_ctx_create failed");
    fl_scene_desc desc;
    fl_gen_describe(&g, &desc);
    fl_scene_t direct = nullptr;
    if (fl_scene_create(ctx, &desc, &direct) != FL_OK) return fail("fl_scene_create failed");
    fl_scene_info info;
    fl_scene_get_info(direct, &info);

    // rays from random points in the scene box in uniform directions
    std::vector<fl_ray> rays(n_rays);
    for (fl_ray &r : rays) {
        r.ox = info.bounds_min[0] + uniform01() * (info.bounds_max[0] - info.bounds_min[0]);
        r.oy = info.bounds_min[1] + uniform01() * (info.bounds_max[1] - info.bounds_min[1]);
        r.oz = info.bounds_min[2] + uniform01() * (info.bounds_max[2] - info.bounds_min[2]);
        const float z = 1.f - 2.f * uniform01(), phi = 6.2831853f * uniform01();
        const float s = std::sqrt(std::fmax(0.f, 1.f - z * z));
        r.dx = s * std::cos(phi);
        r.dy = s * std::sin(phi);
        r.dz = z;
        r.tmin = 1e-4f;
        r.tmax = uniform01() < 0.25f ? 2.f : 1e30f;   // some rays end short of every surface
    }
    std::vector<fl_hit> ref(n_rays);
    if (fl_intersect(direct, rays.data(), ref.data(), n_rays, FL_CLOSEST_HIT, nullptr) != FL_OK)
        return fail("fl_intersect failed");

    // errors: a query before commit, a wrong buffer format
    EmbreeBvh bvh(device);
    {
        RTCRayHit rh;
        std::memset(&rh, 0, sizeof(rh));
        rh.ray.tfar = 1.f;
        rh.hit.geomID = RTC_INVALID_GEOMETRY_ID;
        rtcIntersect1(bvh.scene, &rh);
        if (rtcGetDeviceError(device) != RTC_ERROR_INVALID_OPERATION) return fail("no error for a query before commit");
        RTCGeometry bad = rtcNewGeometry(device, RTC_GEOMETRY_TYPE_TRIANGLE);
        if (rtcSetNewGeometryBuffer(bad, RTC_BUFFER_TYPE_VERTEX, 0, RTC_FORMAT_UINT3, 12, 3) != nullptr ||
            rtcGetDeviceError(device) != RTC_ERROR_INVALID_ARGUMENT)
            return fail("no error for a wrong vertex format");
        rtcReleaseGeometry(bad);
        if (rtcGetDeviceError(device) != RTC_ERROR_NONE) return fail("error not cleared by rtcGetDeviceError");
    }

    // the reference's pattern: one geometry per triangle
    for (uint32_t t = 0; t < nt; ++t) {
        const uint32_t *ix = g.indices + 3 * t;
        Tri tri;
        std::memcpy(&tri.v1, g.vertices + 3 * ix[0], 12);
        std::memcpy(&tri.v2, g.vertices + 3 * ix[1], 12);
        std::memcpy(&tri.v3, g.vertices + 3 * ix[2], 12);
        bvh.add_triangle(tri);
    }
    bvh.finalise();
    if (rtcGetDeviceError(device) != RTC_ERROR_NONE) return fail("scene commit failed");

    uint32_t hits = 0, mismatches = 0, occ_mismatches = 0, ng_mismatches = 0;
    for (uint32_t i = 0; i < n_rays; ++i) {
        RTCRayHit rh;
        const bool hit = bvh.intersect(rays[i], &rh);
        const fl_hit &h = ref[i];
        if (hit != (h.tri != FL_NO_HIT)) { ++mismatches; continue; }
        if (bvh.occluded(rays[i]) != hit) ++occ_mismatches;
        if (!hit) {
            if (rh.ray.tfar != rays[i].tmax) ++mismatches;   // a miss leaves the record untouched
            continue;
        }
        ++hits;
        if (rh.hit.geomID != h.tri || rh.hit.primID != 0 || std::memcmp(&rh.ray.tfar, &h.t, 4) ||
            std::memcmp(&rh.hit.u, &h.u, 4) || std::memcmp(&rh.hit.v, &h.v, 4)) {
            ++mismatches;
            continue;
        }
        const uint32_t *ix = g.indices + 3 * h.tri;
        const Vec3 n = cross_edges(g.vertices + 3 * ix[0], g.vertices + 3 * ix[1], g.vertices + 3 * ix[2]);
        if (!ng_close(n, 1.f, rh.hit)) ++ng_mismatches;
    }

    // meshes of varying size (1, 2, 3, ... triangles each): (geomID, primID) -> triangle
    uint32_t mesh_mismatches = 0, n_meshes = 0;
    {
        EmbreeBvh meshes(device);
        std::vector<uint32_t> first;
        for (uint32_t t = 0, len = 1; t < nt; t += len, ++len) {
            const uint32_t n = len < nt - t ? len : nt - t;
            first.push_back(t);
            meshes.add_mesh(g.vertices, g.indices, t, n);
        }
        meshes.finalise();
        n_meshes = (uint32_t)first.size();
        for (uint32_t i = 0; i < n_rays; ++i) {
            RTCRayHit rh;
            const bool hit = meshes.intersect(rays[i], &rh);
            const fl_hit &h = ref[i];
            if (hit != (h.tri != FL_NO_HIT)) { ++mesh_mismatches; continue; }
            if (!hit) continue;
            if (rh.hit.geomID >= first.size() || first[rh.hit.geomID] + rh.hit.primID != h.tri ||
                std::memcmp(&rh.ray.tfar, &h.t, 4) || std::memcmp(&rh.hit.u, &h.u, 4) ||
                std::memcmp(&rh.hit.v, &h.v, 4))
                ++mesh_mismatches;
        }
        if (rtcGetDeviceError(device) != RTC_ERROR_NONE) return fail("mesh scene failed");
    }

    // opposite winding: one triangle given as (v2, v1, v0), hit at its
    // centroid from the front; Ng is the negated face normal
    {
        EmbreeBvh rev(device);
        const uint32_t t = nt / 2;
        rev.add_mesh(g.vertices, g.indices, t, 1, true);
        rev.finalise();
        const uint32_t *ix = g.indices + 3 * t;
        const float *a = g.vertices + 3 * ix[0], *b = g.vertices + 3 * ix[1], *c = g.vertices + 3 * ix[2];
        const Vec3 n = cross_edges(a, b, c);
        const float len = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
        fl_ray r;
        r.ox = (a[0] + b[0] + c[0]) / 3.f + n.x / len;
        r.oy = (a[1] + b[1] + c[1]) / 3.f + n.y / len;
        r.oz = (a[2] + b[2] + c[2]) / 3.f + n.z / len;
        r.dx = -n.x / len;
        r.dy = -n.y / len;
        r.dz = -n.z / len;
        r.tmin = 0.f;
        r.tmax = 1e30f;
        RTCRayHit rh;
        if (!rev.intersect(r, &rh) || rh.hit.geomID != 0 || rh.hit.primID != 0 || std::fabs(rh.ray.tfar - 1.f) > 1e-3f)
            ++mesh_mismatches;
        else if (!ng_close(n, -1.f, rh.hit))
            ++ng_mismatches;
    }

    fl_scene_destroy(direct);
    fl_ctx_destroy(ctx);
    fl_gen_free(&g);
    rtcReleaseDevice(device);
    const bool ok = mismatches == 0 && occ_mismatches == 0 && ng_mismatches == 0 && mesh_mismatches == 0 && hits > 0;
    printf("{\"ok\": %s, \"triangles\": %u, \"rays\": %u, \"hits\": %u, \"mismatches\": %u, "
           "\"occluded_mismatches\": %u, \"ng_mismatches\": %u, \"meshes\": %u, \"mesh_mismatches\": %u}\n",
           ok ? "true" : "false", nt, n_rays, hits, mismatches, occ_mismatches, ng_mismatches, n_meshes,
           mesh_mismatches);
    return ok ? 0 : 1;
}
