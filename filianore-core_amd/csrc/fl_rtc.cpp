// fl_rtc.cpp -- the Embree 4 API subset of include/fl_rtc.h over the C ABI
// of this library (host code only): geometries collect triangles on the
// host; rtcCommitScene flattens them in attach order into one fl_scene (the
// GPU LBVH); ray queries call fl_intersect.
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "../../include/fl_gpu.h"
#include "../../include/fl_rtc.h"

struct RTCDeviceTy {
    std::atomic<int> refs{1};
    fl_ctx_t ctx = nullptr;
    RTCError err = RTC_ERROR_NONE;
};

struct RTCGeometryTy {
    std::atomic<int> refs{1};
    RTCDeviceTy *dev = nullptr;
    std::vector<float> verts;        // float3
    std::vector<uint32_t> idx;       // uint3
    size_t vstride = 12, istride = 12;
    std::vector<unsigned char> vraw, iraw;   // caller-written buffers (any stride)
    size_t vcount = 0, icount = 0;
    bool valid = true;               // every index addresses a vertex
};

struct RTCSceneTy {
    std::atomic<int> refs{1};
    RTCDeviceTy *dev = nullptr;
    std::vector<RTCGeometryTy *> geoms;
    std::vector<uint32_t> first;     // flattened ID of each geometry's first triangle
    std::vector<float> tri_verts;    // committed: 9 floats per flattened triangle (for Ng)
    fl_scene_t scene = nullptr;
    bool committed = false;
};

namespace {

void set_err(RTCDeviceTy *d, RTCError e) {
    if (d && d->err == RTC_ERROR_NONE) d->err = e;
}

void release_geometry(RTCGeometryTy *g) {
    if (g && --g->refs == 0) delete g;
}

// Looks up the (geomID, primID) of a flattened triangle ID.
void split_id(const RTCSceneTy *s, uint32_t tri, unsigned int *geom, unsigned int *prim) {
    size_t lo = 0, hi = s->first.size();
    while (hi - lo > 1) {
        size_t mid = (lo + hi) / 2;
        if (s->first[mid] <= tri) lo = mid;
        else hi = mid;
    }
    *geom = (unsigned int)lo;
    *prim = tri - s->first[lo];
}

} // namespace

extern "C" {

RTCDevice rtcNewDevice(const char *) {
    RTCDeviceTy *d = new (std::nothrow) RTCDeviceTy();
    if (!d) return nullptr;
    if (fl_ctx_create(0, &d->ctx) != FL_OK) {
        delete d;
        return nullptr;   // no usable GPU
    }
    return d;
}

void rtcRetainDevice(RTCDevice d) {
    if (d) ++d->refs;
}

void rtcReleaseDevice(RTCDevice d) {
    if (d && --d->refs == 0) {
        fl_ctx_destroy(d->ctx);
        delete d;
    }
}

RTCError rtcGetDeviceError(RTCDevice d) {
    if (!d) return RTC_ERROR_INVALID_ARGUMENT;
    RTCError e = d->err;
    d->err = RTC_ERROR_NONE;
    return e;
}

RTCScene rtcNewScene(RTCDevice d) {
    if (!d) return nullptr;
    RTCSceneTy *s = new (std::nothrow) RTCSceneTy();
    if (!s) { set_err(d, RTC_ERROR_OUT_OF_MEMORY); return nullptr; }
    s->dev = d;
    rtcRetainDevice(d);
    return s;
}

void rtcRetainScene(RTCScene s) {
    if (s) ++s->refs;
}

void rtcReleaseScene(RTCScene s) {
    if (!s || --s->refs != 0) return;
    for (RTCGeometryTy *g : s->geoms) release_geometry(g);
    if (s->scene) fl_scene_destroy(s->scene);
    RTCDeviceTy *d = s->dev;
    delete s;
    rtcReleaseDevice(d);
}

RTCGeometry rtcNewGeometry(RTCDevice d, enum RTCGeometryType type) {
    if (!d) return nullptr;
    if (type != RTC_GEOMETRY_TYPE_TRIANGLE) { set_err(d, RTC_ERROR_INVALID_ARGUMENT); return nullptr; }
    RTCGeometryTy *g = new (std::nothrow) RTCGeometryTy();
    if (!g) { set_err(d, RTC_ERROR_OUT_OF_MEMORY); return nullptr; }
    g->dev = d;
    return g;
}

void rtcRetainGeometry(RTCGeometry g) {
    if (g) ++g->refs;
}

void rtcReleaseGeometry(RTCGeometry g) { release_geometry(g); }

void *rtcSetNewGeometryBuffer(RTCGeometry g, enum RTCBufferType type, unsigned int slot, enum RTCFormat format,
                              size_t byteStride, size_t itemCount) {
    if (!g) return nullptr;
    const bool vert = type == RTC_BUFFER_TYPE_VERTEX && format == RTC_FORMAT_FLOAT3;
    const bool index = type == RTC_BUFFER_TYPE_INDEX && format == RTC_FORMAT_UINT3;
    if (slot != 0 || (!vert && !index) || byteStride < 12 || (byteStride & 3)) {
        set_err(g->dev, RTC_ERROR_INVALID_ARGUMENT);
        return nullptr;
    }
    // Embree lets the buffer be read 16 bytes past the last item: pad the same
    std::vector<unsigned char> &raw = vert ? g->vraw : g->iraw;
    raw.assign(byteStride * itemCount + 16, 0);
    (vert ? g->vstride : g->istride) = byteStride;
    (vert ? g->vcount : g->icount) = itemCount;
    return raw.data();
}

void rtcCommitGeometry(RTCGeometry g) {
    if (!g) return;
    g->verts.resize(g->vcount * 3);
    for (size_t i = 0; i < g->vcount; ++i) std::memcpy(&g->verts[3 * i], &g->vraw[i * g->vstride], 12);
    g->idx.resize(g->icount * 3);
    for (size_t i = 0; i < g->icount; ++i) std::memcpy(&g->idx[3 * i], &g->iraw[i * g->istride], 12);
    g->valid = true;
    for (uint32_t v : g->idx)
        if (v >= g->vcount) {
            g->valid = false;
            set_err(g->dev, RTC_ERROR_INVALID_OPERATION);
            break;
        }
}

unsigned int rtcAttachGeometry(RTCScene s, RTCGeometry g) {
    if (!s || !g) return RTC_INVALID_GEOMETRY_ID;
    rtcRetainGeometry(g);
    s->geoms.push_back(g);
    s->committed = false;
    return (unsigned int)(s->geoms.size() - 1);
}

void rtcCommitScene(RTCScene s) {
    if (!s) return;
    if (s->scene) { fl_scene_destroy(s->scene); s->scene = nullptr; }
    std::vector<float> verts;
    std::vector<uint32_t> idx;
    s->first.clear();
    s->tri_verts.clear();
    s->committed = false;
    for (RTCGeometryTy *g : s->geoms)
        if (!g->valid) { set_err(s->dev, RTC_ERROR_INVALID_OPERATION); return; }
    for (RTCGeometryTy *g : s->geoms) {
        s->first.push_back((uint32_t)(idx.size() / 3));
        const uint32_t base = (uint32_t)(verts.size() / 3);
        verts.insert(verts.end(), g->verts.begin(), g->verts.end());
        for (uint32_t v : g->idx) idx.push_back(base + v);
        for (size_t t = 0; t < g->idx.size(); ++t)
            for (int k = 0; k < 3; ++k) s->tri_verts.push_back(g->verts[3 * g->idx[t] + k]);
    }
    if (idx.empty()) {   // an empty scene is valid: every ray misses
        s->committed = true;
        return;
    }
    fl_material m;
    std::memset(&m, 0, sizeof(m));
    m.kind = FL_MAT_LAMBERT;
    m.rgb[0] = m.rgb[1] = m.rgb[2] = 0.5f;
    m.weight = 1.f;
    fl_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.vertices = verts.data();
    d.n_vertices = (uint32_t)(verts.size() / 3);
    d.indices = idx.data();
    d.n_triangles = (uint32_t)(idx.size() / 3);
    d.materials = &m;
    d.n_materials = 1;
    if (fl_scene_create(s->dev->ctx, &d, &s->scene) != FL_OK) {
        s->scene = nullptr;
        set_err(s->dev, RTC_ERROR_UNKNOWN);
        return;
    }
    s->committed = true;
}

void rtcInitRayQueryContext(struct RTCRayQueryContext *c) {
    if (c) c->instID[0] = RTC_INVALID_GEOMETRY_ID;
}

void rtcIntersect1(RTCScene s, struct RTCRayHit *rh, struct RTCIntersectArguments *) {
    if (!s || !rh) return;
    if (!s->committed) { set_err(s->dev, RTC_ERROR_INVALID_OPERATION); return; }
    if (!s->scene) return;
    fl_ray r{rh->ray.org_x, rh->ray.org_y, rh->ray.org_z, rh->ray.tnear,
             rh->ray.dir_x, rh->ray.dir_y, rh->ray.dir_z, rh->ray.tfar};
    fl_hit h;
    if (fl_intersect(s->scene, &r, &h, 1, FL_CLOSEST_HIT, nullptr) != FL_OK) {
        set_err(s->dev, RTC_ERROR_UNKNOWN);
        return;
    }
    if (h.tri == FL_NO_HIT) return;   // rayhit untouched: geomID keeps RTC_INVALID_GEOMETRY_ID
    const float *v = &s->tri_verts[(size_t)h.tri * 9];
    const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    rh->hit.Ng_x = e1[1] * e2[2] - e1[2] * e2[1];
    rh->hit.Ng_y = e1[2] * e2[0] - e1[0] * e2[2];
    rh->hit.Ng_z = e1[0] * e2[1] - e1[1] * e2[0];
    rh->hit.u = h.u;
    rh->hit.v = h.v;
    split_id(s, h.tri, &rh->hit.geomID, &rh->hit.primID);
    rh->hit.instID[0] = RTC_INVALID_GEOMETRY_ID;
    rh->ray.tfar = h.t;
}

void rtcOccluded1(RTCScene s, struct RTCRay *ray, struct RTCOccludedArguments *) {
    if (!s || !ray) return;
    if (!s->committed) { set_err(s->dev, RTC_ERROR_INVALID_OPERATION); return; }
    if (!s->scene) return;
    fl_ray r{ray->org_x, ray->org_y, ray->org_z, ray->tnear, ray->dir_x, ray->dir_y, ray->dir_z, ray->tfar};
    fl_hit h;
    if (fl_intersect(s->scene, &r, &h, 1, FL_ANY_HIT, nullptr) != FL_OK) {
        set_err(s->dev, RTC_ERROR_UNKNOWN);
        return;
    }
    if (h.tri != FL_NO_HIT) ray->tfar = -std::numeric_limits<float>::infinity();
}

} // extern "C"
