// fl_trace.hip -- closest/any-hit queries over ray arrays (Scene::intersect,
// Scene::intersect_p: core/scene.cpp:25-31, bvh/bvh.cpp:49-88).
//
// Persistent grid sized by occupancy; lanes refill with new rays as their ray
// finishes, so the global spill area of the short stack is bounded by the
// grid, not by n.
#include <mutex>
#include <vector>

#include "fl_internal.h"
#include "fl_trace.h"

// The fixed scratch of fl_intersect_device_noalloc (refill cursor, then the
// spill area of the traversal stacks): every call on a device shares that
// device's copy, so calls that overlap on one device race.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kBlock = 256;
constexpr int kStack = 8;
constexpr int kBudget = 24;

struct ApiSrc {
    static constexpr bool kMulti = false;
    const fl_ray *rays;
    fl_hit *hits;
    bool any;
    fl::TraceScene sc;
    __device__ fl::TraceScene base_scene() const { return sc; }
    __device__ void load(uint32_t i, fl::v3 &o, fl::v3 &d, float &tmin, float &tmax, uint32_t &anyw,
                         fl::TraceScene &ray_sc) const {
        anyw = any ? 0u : 0x80000000u;
        ray_sc = sc;
        const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
        float4 a = rp[0], b = rp[1];
        o = fl::v3(a.x, a.y, a.z);
        d = fl::v3(b.x, b.y, b.z);
        tmin = a.w;
        tmax = b.w;
    }
    __device__ void store(uint32_t i, const fl::TravState &s) const {
        bool hit = s.best_id != fl::kNoHit;
        float4 out;
        out.x = hit ? s.best_t : __uint_as_float(0x7f800000u);
        out.y = __uint_as_float(hit ? s.best_id : fl::kNoHit);
        out.z = hit ? s.bu : 0.f;
        out.w = hit ? s.bv : 0.f;
        reinterpret_cast<float4 *>(hits)[i] = out;
    }
};

__global__ void k_zero_fetch(uint32_t *p) { *p = 0u; }

template <bool ANY, bool SKIP, bool COUNT, bool PREFETCH>
__global__ void __launch_bounds__(kBlock) k_trace_api(fl::TraceScene sc, ApiSrc src, uint32_t n, uint2 *spill,
                                                      uint32_t *fetch, unsigned long long *counters) {
    __shared__ unsigned long long stk[kStack * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    fl::LaneStack ls;
    ls.lds = (fl::lds_u64 *)(stk + threadIdx.x);
    ls.lds_stride = kBlock;
    ls.cap = kStack;
    ls.spill = spill + gtid;
    ls.spill_stride = gridDim.x * kBlock;
    fl::TravCounters c;
    fl::trace_persistent<COUNT, SKIP, PREFETCH>(src, n, fetch, ls, kBudget, 16, c);
    if (COUNT) {
        atomicAdd(&counters[0], (unsigned long long)c.nodes);
        atomicAdd(&counters[1], (unsigned long long)c.tris);
        atomicAdd(&counters[2], (unsigned long long)c.spills);
    }
}

// Where a launch keeps its refill cursor (256 bytes) and its spill area:
// the context's grown buffer, or a fixed area whose size bounds the grid.
struct TraceScratch {
    void *base;          // nullptr: grow ctx->scratch as needed
    size_t bytes;        // capacity of a fixed area
    hipStream_t stream;
};

template <bool ANY, bool SKIP, bool COUNT, bool PREFETCH>
int launch(fl_ctx_t ctx, int depth, fl::TraceScene sc, ApiSrc src, uint32_t n, unsigned long long *c,
           const TraceScratch &ts) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, reinterpret_cast<const void *>(k_trace_api<ANY, SKIP, COUNT, PREFETCH>), kBlock, 0) != hipSuccess ||
        occ <= 0)
        occ = 4;
    uint32_t grid = (uint32_t)(ctx->n_cu * occ);
    uint32_t need = (n + kBlock - 1) / kBlock;
    if (grid > need) grid = need;
    const size_t per_lane = sizeof(uint2) * (size_t)(depth > kStack ? depth - kStack + 1 : 1);
    void *base = ts.base;
    if (base) {
        // a fixed area: fewer persistent blocks when the BVH is deep (each
        // lane still takes rays until none are left)
        const size_t fit = ts.bytes > 256 ? (ts.bytes - 256) / (per_lane * kBlock) : 0;
        if (fit == 0)
            return flx::set_error(ctx, FL_E_OVERFLOW, "BVH too deep for the fixed traversal scratch");
        if (grid > fit) grid = (uint32_t)fit;
    } else {
        const size_t bytes = (size_t)grid * kBlock * per_lane + 256;
        if (ctx->scratch_bytes < bytes) {
            if (ctx->scratch) (void)hipFree(ctx->scratch);
            ctx->scratch = nullptr;
            ctx->scratch_bytes = 0;
            if (hipMalloc(&ctx->scratch, bytes) != hipSuccess) return flx::set_error(ctx, FL_E_OOM, "spill stack");
            ctx->scratch_bytes = bytes;
        }
        base = ctx->scratch;
    }
    uint32_t *fetch = reinterpret_cast<uint32_t *>(base);
    uint2 *spill = reinterpret_cast<uint2 *>(reinterpret_cast<char *>(base) + 256);
    // the cursor starts at zero: a kernel, not a memset, so that the call can
    // be captured into a graph and replayed
    hipLaunchKernelGGL(k_zero_fetch, dim3(1), dim3(1), 0, ts.stream, fetch);
    hipLaunchKernelGGL((k_trace_api<ANY, SKIP, COUNT, PREFETCH>), dim3(grid), dim3(kBlock), 0, ts.stream, sc, src, n,
                       spill, fetch, c);
    return FL_OK;
}

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;   // per device: the device address of `scratch`

} // namespace

namespace flx {

int noalloc_scratch_init(fl_ctx_t ctx) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.size() <= (size_t)ctx->device) g_scratch_ptr.resize((size_t)ctx->device + 1, nullptr);
    if (!g_scratch_ptr[ctx->device]) {
        // the device current (fl_ctx_create set it): each device holds its own
        // copy of the array at its own address; host code uses only that
        // address, never the array's name
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return set_error(ctx, FL_E_HIP, "traversal scratch: hipGetSymbolAddress failed");
        g_scratch_ptr[ctx->device] = p;
    }
    ctx->noalloc_scratch = g_scratch_ptr[ctx->device];
    return FL_OK;
}

size_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

int trace_rays(fl_ctx_t ctx, fl_scene_s *s, const fl_ray *d_rays, fl_hit *d_hits, uint32_t n, uint32_t flags,
               unsigned long long *d_counters, bool fixed_scratch, hipStream_t stream) {
    if (n == 0) return FL_OK;
    TraceScratch ts{nullptr, 0, stream ? stream : ctx->stream};
    if (fixed_scratch) {
        if (!ctx->noalloc_scratch) return set_error(ctx, FL_E_INVALID, "traversal scratch not set up");
        ts.base = ctx->noalloc_scratch;
        ts.bytes = kNoAllocScratchBytes;
    }
    fl::TraceScene sc{s->d_nodes, s->d_tris};
    ApiSrc src{d_rays, d_hits, (flags & FL_ANY_HIT) != 0, sc};
    const int depth = (int)s->depth;
    bool any = flags & FL_ANY_HIT, skip = flags & FL_SKIP_LIGHTS, cnt = (flags & FL_COUNTERS) && d_counters;
    // the render path's default refill mode; counting runs use the refill mode
    const bool pf = s->n_tris >= fl::kPrefetchTris;
    int rc;
    if (cnt) {
        if (!any) rc = launch<false, false, true, false>(ctx, depth, sc, src, n, d_counters, ts);
        else if (skip) rc = launch<true, true, true, false>(ctx, depth, sc, src, n, d_counters, ts);
        else rc = launch<true, false, true, false>(ctx, depth, sc, src, n, d_counters, ts);
    } else if (pf) {
        if (!any) rc = launch<false, false, false, true>(ctx, depth, sc, src, n, d_counters, ts);
        else if (skip) rc = launch<true, true, false, true>(ctx, depth, sc, src, n, d_counters, ts);
        else rc = launch<true, false, false, true>(ctx, depth, sc, src, n, d_counters, ts);
    } else {
        if (!any) rc = launch<false, false, false, false>(ctx, depth, sc, src, n, d_counters, ts);
        else if (skip) rc = launch<true, true, false, false>(ctx, depth, sc, src, n, d_counters, ts);
        else rc = launch<true, false, false, false>(ctx, depth, sc, src, n, d_counters, ts);
    }
    if (rc) return rc;
    FL_HIP_CHECK(ctx, hipGetLastError());
    return FL_OK;
}

} // namespace flx
