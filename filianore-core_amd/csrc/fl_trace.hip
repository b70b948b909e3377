// fl_trace.hip -- closest/any-hit queries over ray arrays (Scene::intersect,
// Scene::intersect_p: core/scene.cpp:25-31, bvh/bvh.cpp:49-88).
//
// Persistent grid: each lane loops over rays with a grid stride, so the
// global spill area of the short stack is bounded by the grid, not by n.
#include "fl_internal.h"
#include "fl_trace.h"

namespace {

constexpr int kBlock = 256;

template <bool ANY, bool SKIP, bool COUNT, int S>
__global__ void __launch_bounds__(kBlock) k_trace_api(fl::TraceScene sc, const fl_ray *__restrict__ rays,
                                                      fl_hit *__restrict__ hits, uint32_t n, uint2 *spill,
                                                      unsigned long long *counters) {
    __shared__ uint2 stk[S * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t nthreads = gridDim.x * kBlock;
    fl::LaneStack ls;
    ls.lds = stk + threadIdx.x;
    ls.lds_stride = kBlock;
    ls.cap = S;
    ls.spill = spill + gtid;
    ls.spill_stride = nthreads;
    uint32_t cn = 0, ct = 0, cs = 0;
    for (uint32_t i = gtid; i < n; i += nthreads) {
        const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
        float4 a = rp[0], b = rp[1];
        fl::TraceHit h;
        bool hit = fl::traverse<ANY, SKIP, COUNT>(sc, fl::v3(a.x, a.y, a.z), fl::v3(b.x, b.y, b.z), a.w, b.w, ls, h,
                                                   cn, ct, cs);
        float4 out;
        out.x = hit ? h.t : __uint_as_float(0x7f800000u);
        out.y = __uint_as_float(hit ? h.id : fl::kNoHit);
        out.z = hit ? h.u : 0.f;
        out.w = hit ? h.v : 0.f;
        reinterpret_cast<float4 *>(hits)[i] = out;
    }
    if (COUNT) {
        atomicAdd(&counters[0], (unsigned long long)cn);
        atomicAdd(&counters[1], (unsigned long long)ct);
        atomicAdd(&counters[2], (unsigned long long)cs);
    }
}

template <bool ANY, bool SKIP, bool COUNT>
void launch(hipStream_t st, int grid, fl::TraceScene sc, const fl_ray *r, fl_hit *h, uint32_t n, uint2 *spill,
            unsigned long long *c) {
    hipLaunchKernelGGL((k_trace_api<ANY, SKIP, COUNT, 16>), dim3(grid), dim3(kBlock), 0, st, sc, r, h, n, spill, c);
}

} // namespace

namespace flx {

int trace_rays(fl_ctx_t ctx, fl_scene_s *s, const fl_ray *d_rays, fl_hit *d_hits, uint32_t n, uint32_t flags,
               unsigned long long *d_counters) {
    if (n == 0) return FL_OK;
    const int S = 16;
    int grid = ctx->n_cu * 8;
    uint32_t need = (uint32_t)((n + kBlock - 1) / kBlock);
    if ((uint32_t)grid > need) grid = (int)need;
    size_t spill_entries = (size_t)grid * kBlock * (size_t)(s->depth > (uint32_t)S ? s->depth - S + 1 : 1);
    size_t bytes = spill_entries * sizeof(uint2);
    if (ctx->scratch_bytes < bytes) {
        if (ctx->scratch) (void)hipFree(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        if (hipMalloc(&ctx->scratch, bytes) != hipSuccess) return set_error(ctx, FL_E_OOM, "spill stack");
        ctx->scratch_bytes = bytes;
    }
    fl::TraceScene sc{s->d_nodes, s->d_tris};
    uint2 *spill = (uint2 *)ctx->scratch;
    bool any = flags & FL_ANY_HIT, skip = flags & FL_SKIP_LIGHTS, cnt = (flags & FL_COUNTERS) && d_counters;
    hipStream_t st = ctx->stream;
    if (!any) {
        if (cnt) launch<false, false, true>(st, grid, sc, d_rays, d_hits, n, spill, d_counters);
        else launch<false, false, false>(st, grid, sc, d_rays, d_hits, n, spill, d_counters);
    } else if (skip) {
        if (cnt) launch<true, true, true>(st, grid, sc, d_rays, d_hits, n, spill, d_counters);
        else launch<true, true, false>(st, grid, sc, d_rays, d_hits, n, spill, d_counters);
    } else {
        if (cnt) launch<true, false, true>(st, grid, sc, d_rays, d_hits, n, spill, d_counters);
        else launch<true, false, false>(st, grid, sc, d_rays, d_hits, n, spill, d_counters);
    }
    FL_HIP_CHECK(ctx, hipGetLastError());
    return FL_OK;
}

} // namespace flx
