// Ray-batch traversal kernels behind lmgpu_trace_closest / lmgpu_trace_any / lmgpu_trace_count.
#include "context.hpp"
#include "traverse.hpp"

namespace lmgpu {

struct RaysSoA {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
};

__device__ __forceinline__ RayIn load_ray(const RaysSoA& r, uint32_t i) {
    RayIn x;
    x.ox = r.ox[i]; x.oy = r.oy[i]; x.oz = r.oz[i];
    x.dx = r.dx[i]; x.dy = r.dy[i]; x.dz = r.dz[i];
    x.tmin = r.tmin[i]; x.tmax = r.tmax[i];
    return x;
}

template <bool kRefine>
__global__ __launch_bounds__(kTraceBlock) void k_trace_closest(const Node4* __restrict__ nodes,
                                                               const TriRecord* __restrict__ tris, RaysSoA rays,
                                                               uint32_t n, float* __restrict__ t_out,
                                                               float* __restrict__ b0_out, float* __restrict__ b1_out,
                                                               int32_t* __restrict__ prim_out) {
    __shared__ uint32_t lds_stack[kStackLDS * kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const RayIn r = load_ray(rays, i);
    Stack st;
    st.lds = lds_stack;
    st.lane = int(threadIdx.x);
    st.overflow = false;
    HitOut h;
    TraceCounters cnt{0, 0};
    if (traverse<false, false>(nodes, tris, r, st, h, cnt)) {
        float t, b0, b1;
        winner_barycentrics(tris, r, h, kRefine, t, b0, b1);
        t_out[i] = t;
        b0_out[i] = b0;
        b1_out[i] = b1;
        prim_out[i] = int32_t(h.prim);
    } else {
        t_out[i] = -1.f;
        b0_out[i] = 0.f;
        b1_out[i] = 0.f;
        prim_out[i] = -1;
    }
}

__global__ __launch_bounds__(kTraceBlock) void k_trace_any(const Node4* __restrict__ nodes,
                                                           const TriRecord* __restrict__ tris, RaysSoA rays,
                                                           uint32_t n, uint8_t* __restrict__ occluded) {
    __shared__ uint32_t lds_stack[kStackLDS * kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const RayIn r = load_ray(rays, i);
    Stack st;
    st.lds = lds_stack;
    st.lane = int(threadIdx.x);
    st.overflow = false;
    HitOut h;
    TraceCounters cnt{0, 0};
    occluded[i] = traverse<true, false>(nodes, tris, r, st, h, cnt) ? 1 : 0;
}

__global__ __launch_bounds__(kTraceBlock) void k_trace_count(const Node4* __restrict__ nodes,
                                                             const TriRecord* __restrict__ tris, RaysSoA rays,
                                                             uint32_t n, unsigned long long* __restrict__ counts) {
    __shared__ uint32_t lds_stack[kStackLDS * kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    TraceCounters cnt{0, 0};
    if (i < n) {
        const RayIn r = load_ray(rays, i);
        Stack st;
        st.lds = lds_stack;
        st.lane = int(threadIdx.x);
        st.overflow = false;
        HitOut h;
        traverse<false, true>(nodes, tris, r, st, h, cnt);
    }
    // Wave-level reduction, one atomic per wave per counter.
    unsigned long long a = cnt.nodes, b = cnt.tris;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&counts[0], a);
        atomicAdd(&counts[1], b);
    }
}

static RaysSoA soa_of(const float* const* r) { return RaysSoA{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]}; }

hipError_t launch_trace_closest(lmgpu_ctx* ctx, const float* const* rays, uint32_t n, float* t, float* b0, float* b1,
                                int32_t* prim, bool refine) {
    const dim3 grid((n + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
    return timed(ctx, kKTraceClosest, [&]() {
        if (refine)
            hipLaunchKernelGGL(k_trace_closest<true>, grid, block, 0, ctx->stream, ctx->nodes.ptr, ctx->tris.ptr,
                               soa_of(rays), n, t, b0, b1, prim);
        else
            hipLaunchKernelGGL(k_trace_closest<false>, grid, block, 0, ctx->stream, ctx->nodes.ptr, ctx->tris.ptr,
                               soa_of(rays), n, t, b0, b1, prim);
    });
}

hipError_t launch_trace_any(lmgpu_ctx* ctx, const float* const* rays, uint32_t n, uint8_t* occluded) {
    const dim3 grid((n + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
    return timed(ctx, kKTraceAny, [&]() {
        hipLaunchKernelGGL(k_trace_any, grid, block, 0, ctx->stream, ctx->nodes.ptr, ctx->tris.ptr, soa_of(rays), n,
                           occluded);
    });
}

hipError_t launch_trace_count(lmgpu_ctx* ctx, const float* const* rays, uint32_t n, unsigned long long* counts) {
    const dim3 grid((n + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
    hipLaunchKernelGGL(k_trace_count, grid, block, 0, ctx->stream, ctx->nodes.ptr, ctx->tris.ptr, soa_of(rays), n,
                       counts);
    return hipGetLastError();
}

}  // namespace lmgpu
