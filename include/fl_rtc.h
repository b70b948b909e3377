/*
 * fl_rtc.h -- the subset of Embree 4's public C API that filianore-core's
 * BVH uses (bvh/bvh.h:21-45, bvh/bvh.cpp:7-88), implemented over
 * libfl_gpu.so: source written against Embree for this call pattern compiles
 * against this header and traverses the GPU LBVH.
 *
 * Names, argument meanings and the ray/hit record fields follow Embree 4's
 * published API (rtcore_device.h, rtcore_scene.h, rtcore_geometry.h,
 * rtcore_ray.h); the code behind them is this library's:
 *   rtcNewDevice / rtcReleaseDevice      -> fl_ctx_create / fl_ctx_destroy (device 0)
 *   rtcNewScene, rtcNewGeometry (triangles only), rtcSetNewGeometryBuffer
 *     (float3 vertices, uint3 indices), rtcCommitGeometry, rtcAttachGeometry,
 *     rtcReleaseGeometry, rtcCommitScene -> one fl_scene over all attached
 *     geometries in attach order (geomID, primID map to the flattened
 *     triangle ID), built by fl_scene_create
 *   rtcIntersect1 -> fl_intersect closest hit (t in (tnear, tfar], watertight)
 *   rtcOccluded1  -> fl_intersect any hit; an occluded ray gets tfar = -inf
 * Ng is the unnormalized (v1 - v0) x (v2 - v0); u, v weight v1 and v2.
 * Differences from Embree: errors are reported through rtcGetDeviceError as
 * RTC_ERROR_UNKNOWN / RTC_ERROR_INVALID_ARGUMENT / ..., not through an
 * error callback; one ray per call is exact but launches one traversal,
 * so batch callers should use fl_intersect directly. No counterpart of the
 * reference's own code: this replaces the Embree library it links.
 */
#ifndef FL_RTC_H
#define FL_RTC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTC_INVALID_GEOMETRY_ID ((unsigned int)-1)
#define RTC_MAX_INSTANCE_LEVEL_COUNT 1

typedef struct RTCDeviceTy *RTCDevice;
typedef struct RTCSceneTy *RTCScene;
typedef struct RTCGeometryTy *RTCGeometry;

enum RTCError {
    RTC_ERROR_NONE = 0,
    RTC_ERROR_UNKNOWN = 1,
    RTC_ERROR_INVALID_ARGUMENT = 2,
    RTC_ERROR_INVALID_OPERATION = 3,
    RTC_ERROR_OUT_OF_MEMORY = 4,
    RTC_ERROR_UNSUPPORTED_CPU = 5,
    RTC_ERROR_CANCELLED = 6
};

enum RTCGeometryType { RTC_GEOMETRY_TYPE_TRIANGLE = 0 };
enum RTCBufferType { RTC_BUFFER_TYPE_INDEX = 0, RTC_BUFFER_TYPE_VERTEX = 1 };
enum RTCFormat { RTC_FORMAT_UNDEFINED = 0, RTC_FORMAT_UINT3 = 0x5003, RTC_FORMAT_FLOAT3 = 0x9003 };

struct RTCRay {
    float org_x, org_y, org_z;
    float tnear;
    float dir_x, dir_y, dir_z;
    float time;
    float tfar;
    unsigned int mask;
    unsigned int id;
    unsigned int flags;
};

struct RTCHit {
    float Ng_x, Ng_y, Ng_z;
    float u, v;
    unsigned int primID;
    unsigned int geomID;
    unsigned int instID[RTC_MAX_INSTANCE_LEVEL_COUNT];
};

struct RTCRayHit {
    struct RTCRay ray;
    struct RTCHit hit;
};

struct RTCRayQueryContext {
    unsigned int instID[RTC_MAX_INSTANCE_LEVEL_COUNT];
};

struct RTCIntersectArguments;
struct RTCOccludedArguments;

/* the optional trailing arguments of the ray queries default to NULL in C++ */
#ifdef __cplusplus
#define FL_RTC_OPTIONAL = NULL
#else
#define FL_RTC_OPTIONAL
#endif

RTCDevice rtcNewDevice(const char *config);
void rtcRetainDevice(RTCDevice device);
void rtcReleaseDevice(RTCDevice device);
enum RTCError rtcGetDeviceError(RTCDevice device);

RTCScene rtcNewScene(RTCDevice device);
void rtcRetainScene(RTCScene scene);
void rtcReleaseScene(RTCScene scene);
void rtcCommitScene(RTCScene scene);

RTCGeometry rtcNewGeometry(RTCDevice device, enum RTCGeometryType type);
void rtcRetainGeometry(RTCGeometry geometry);
void rtcReleaseGeometry(RTCGeometry geometry);
void *rtcSetNewGeometryBuffer(RTCGeometry geometry, enum RTCBufferType type, unsigned int slot, enum RTCFormat format,
                              size_t byteStride, size_t itemCount);
void rtcCommitGeometry(RTCGeometry geometry);
unsigned int rtcAttachGeometry(RTCScene scene, RTCGeometry geometry);

void rtcInitRayQueryContext(struct RTCRayQueryContext *context);
void rtcIntersect1(RTCScene scene, struct RTCRayHit *rayhit, struct RTCIntersectArguments *args FL_RTC_OPTIONAL);
void rtcOccluded1(RTCScene scene, struct RTCRay *ray, struct RTCOccludedArguments *args FL_RTC_OPTIONAL);

#ifdef __cplusplus
}
#endif
#endif /* FL_RTC_H */
