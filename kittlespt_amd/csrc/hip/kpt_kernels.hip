// kpt_kernels.hip — traversal and megakernel path-tracing kernels for gfx950,
// with their host launch wrappers.
#include "kpt_integrator.h"
#include "kpt_launch.h"
#include "kpt_temporal.h"

namespace kpt {

// ------------------------------------------------------------------ traversal
template <bool COUNT>
__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_closest(SceneView S, const kpt_ray* __restrict__ rays,
                                                               kpt_hit* __restrict__ hits, int64_t n,
                                                               int32_t* overflow,
                                                               unsigned long long* counters) {
    __shared__ uint32_t s_stack[LDS_STACK][TRACE_BLOCK];
    const int64_t i = (int64_t)blockIdx.x * TRACE_BLOCK + threadIdx.x;
    TraceCounts c = {0, 0, 0, 0, 0};
    if (i < n) {
        const kpt_ray r = rays[i];
        Ray w = make_ray(mk3(r.org[0], r.org[1], r.org[2]), mk3(r.dir[0], r.dir[1], r.dir[2]));
        bool ovf = false;
        Hit h = trace_closest<COUNT>(S, w, r.tmin, r.tmax, StackShape{&s_stack[0][threadIdx.x], TRACE_BLOCK, LDS_STACK}, &ovf, &c);
        kpt_hit o;
        o.t = h.t;
        if (h.slot >= 0) {
            o.prim_id = S.tris[h.slot].prim;
            o.inst_id = h.inst;
            o.bary[0] = 1.0f - h.u - h.v;
            o.bary[1] = h.u;
            o.bary[2] = h.v;
        } else {
            o.prim_id = -1; o.inst_id = -1; o.bary[0] = 0.0f; o.bary[1] = 0.0f; o.bary[2] = 0.0f;
        }
        hits[i] = o;
        if (ovf) atomicOr(overflow, 1);
    }
    if (COUNT) {
        // wave-level reduction, then one atomic per wave and counter
        uint32_t v[5] = {c.interior, c.leaf, c.tri, c.entries, c.tlas};
        for (int k = 0; k < 5; ++k) {
            unsigned long long x = v[k];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
            if ((threadIdx.x & 63) == 0) atomicAdd(&counters[k], x);
        }
    }
}

__global__ __launch_bounds__(TRACE_BLOCK) void k_trace_any(SceneView S, const kpt_ray* __restrict__ rays,
                                                           int32_t* __restrict__ occluded, int64_t n,
                                                           int32_t* overflow) {
    __shared__ uint32_t s_stack[LDS_STACK][TRACE_BLOCK];
    const int64_t i = (int64_t)blockIdx.x * TRACE_BLOCK + threadIdx.x;
    if (i >= n) return;
    const kpt_ray r = rays[i];
    Ray w = make_ray(mk3(r.org[0], r.org[1], r.org[2]), mk3(r.dir[0], r.dir[1], r.dir[2]));
    bool ovf = false;
    bool o = trace_any<false>(S, w, r.tmin, r.tmax, StackShape{&s_stack[0][threadIdx.x], TRACE_BLOCK, LDS_STACK}, &ovf, nullptr);
    occluded[i] = o ? 1 : 0;
    if (ovf) atomicOr(overflow, 1);
}

hipError_t launch_trace_closest(hipStream_t s, const SceneView& S, const kpt_ray* rays, kpt_hit* hits, int64_t n,
                                int32_t* overflow, kpt_trace_counters* counters) {
    if (n <= 0) return hipSuccess;
    dim3 grid((unsigned)((n + TRACE_BLOCK - 1) / TRACE_BLOCK));
    if (counters) {
        hipLaunchKernelGGL(k_trace_closest<true>, grid, dim3(TRACE_BLOCK), 0, s, S, rays, hits, n, overflow,
                           reinterpret_cast<unsigned long long*>(counters));
    } else {
        hipLaunchKernelGGL(k_trace_closest<false>, grid, dim3(TRACE_BLOCK), 0, s, S, rays, hits, n, overflow,
                           (unsigned long long*)nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_trace_any(hipStream_t s, const SceneView& S, const kpt_ray* rays, int32_t* occluded, int64_t n,
                            int32_t* overflow) {
    if (n <= 0) return hipSuccess;
    dim3 grid((unsigned)((n + TRACE_BLOCK - 1) / TRACE_BLOCK));
    hipLaunchKernelGGL(k_trace_any, grid, dim3(TRACE_BLOCK), 0, s, S, rays, occluded, n, overflow);
    return hipGetLastError();
}

// ------------------------------------------------------------------ megakernel
// computePathTraceSamplesMegaKernel (kernels.cu:385-451) with static
// accumulation (Integrator::addSample, integrator.cuh:406-416) or temporal
// reprojection (reprojectAccumulate, kernels.cu:299-338).  16x16 pixel
// tiles, XCD-swizzled tile order.
constexpr int TILE = 16;

__global__ __launch_bounds__(TRACE_BLOCK) void k_megakernel(SceneView S, FrameConst fc, int32_t frame,
                                                            FrameBuffers fb, int32_t* overflow) {
    __shared__ uint32_t s_stack[LDS_STACK][TRACE_BLOCK];
    const uint32_t tiles_x = (fc.width + TILE - 1) / TILE;
    const uint32_t tiles_y = (fc.height + TILE - 1) / TILE;
    const uint32_t tile = xcd_swizzle(blockIdx.x, tiles_x * tiles_y);
    const int px = (int)((tile % tiles_x) * TILE + (threadIdx.x % TILE));
    const int py = (int)((tile / tiles_x) * TILE + (threadIdx.x / TILE));
    if (px >= fc.width || py >= fc.height) return;
    const int64_t idx = (int64_t)py * fc.width + px;
    uint32_t seed = init_seed(px, py, fc.width, frame, fc.seed_mode);
    // BoxFilter::sample consumes two numbers whose jitter is discarded (kernels.cu:397-402)
    (void)rng1(seed);
    (void)rng1(seed);
    Ray ray = camera_ray(fc, px, py);
    VisOut vs;
    vs.albedo = mk3(0.0f, 0.0f, 0.0f); vs.normal = mk3(0.0f, 0.0f, 0.0f); vs.pos = mk3(0.0f, 0.0f, 0.0f);
    vs.depth = __builtin_inff(); vs.inst = -1; vs.prim = -1;
    bool ovf = false;
    v3 L = li_path(S, fc, ray, seed, StackShape{&s_stack[0][threadIdx.x], TRACE_BLOCK, LDS_STACK}, &ovf, &vs);
    L = (mk3(1.0f, 1.0f, 1.0f) * 1.0f) * L;
    L = clamp_output(L);
    write_pixel(S, fc, fb, frame, idx, px, py, L, vs.albedo, vs.normal, vs.depth, vs.inst, vs.prim, vs.pos);
    if (ovf) atomicOr(overflow, 1);
}

hipError_t launch_megakernel(hipStream_t s, const SceneView& S, const FrameConst& fc, int32_t frame,
                             const FrameBuffers& fb, int32_t* overflow) {
    const uint32_t tiles = ((fc.width + TILE - 1) / TILE) * ((fc.height + TILE - 1) / TILE);
    hipLaunchKernelGGL(k_megakernel, dim3(tiles), dim3(TRACE_BLOCK), 0, s, S, fc, frame, fb, overflow);
    return hipGetLastError();
}

__global__ void k_clear4(float4* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

hipError_t launch_clear(hipStream_t s, float4* p, int64_t n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace kpt
