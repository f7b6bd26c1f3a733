// kpt_wavefront.hip — wavefront form of the path tracer for gfx950.
//
// The megakernel runs Integrator::Li (integrator.cuh:241-402) to completion in
// one thread.  Here each bounce is a sequence of stage kernels over queues of
// live paths whose state stays in HBM (struct of arrays):
//   generate -> [ trace closest -> sort by material -> shade -> trace shadow ]*
//   -> finish (clamp, G/V-buffer, accumulation)
// Per-path random-number consumption and the order of additions to L are the
// megakernel's (the shadow stage of bounce k adds its term before bounce k+1
// shades), so the radiance of every pixel is bit-identical to the megakernel.
// Trace and shade stages are persistent: a grid sized to the device pulls
// 256-path chunks from a per-stage atomic counter until the queue is drained.
#include "kpt_integrator.h"
#include "kpt_launch.h"

namespace kpt {

namespace {

constexpr int WF_BLOCK = 256;
constexpr int MAX_KEYS = 64;
// counters[] slots (zeroed once per frame by a fill kernel)
constexpr int C_QCOUNT = 0;      // + bounce parity: [0],[1] live-queue sizes
constexpr int C_SHADOW = 2;      // shadow-queue size of the current bounce
constexpr int C_PULL = 4;        // 4..(4+3*MAX_BOUNCE): pull counters per stage and bounce
constexpr int MAX_BOUNCE_SLOTS = 40;

KD uint32_t fbits(float f) { return __builtin_bit_cast(uint32_t, f); }
KD float bitsf(uint32_t u) { return __builtin_bit_cast(float, u); }

// Grabs the next chunk of `count` for this workgroup; uniform across the block.
__device__ __forceinline__ int pull_chunk(int32_t* ctr, int* s_base) {
    if (threadIdx.x == 0) *s_base = atomicAdd(ctr, WF_BLOCK);
    __syncthreads();
    int b = *s_base;
    __syncthreads();
    return b;
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Camera ray + seed for every pixel (kernels.cu:385-407).
__global__ __launch_bounds__(WF_BLOCK) void k_generate(FrameConst fc, int32_t frame, WavefrontState W) {
    const int64_t npx = (int64_t)fc.width * fc.height;
    const int64_t i = (int64_t)blockIdx.x * WF_BLOCK + threadIdx.x;
    if (i >= npx) return;
    const int px = (int)(i % fc.width), py = (int)(i / fc.width);
    uint32_t seed = init_seed(px, py, fc.width, frame, fc.seed_mode);
    (void)rng1(seed);
    (void)rng1(seed);
    Ray r = camera_ray(fc, px, py);
    W.ray_o[i] = make_float4(r.o.x, r.o.y, r.o.z, bitsf((uint32_t)i));
    W.ray_d[i] = make_float4(r.d.x, r.d.y, r.d.z, bitsf(seed));
    W.beta[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    W.L[i] = make_float4(0.0f, 0.0f, 0.0f, bitsf(0u));
    W.prev_pos[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    W.queue[0][i] = (int32_t)i;
    if (i == 0) W.counters[C_QCOUNT] = (int32_t)npx;
}

// Closest hit of every live path.  COUNT accumulates traversal events.
template <bool COUNT>
__global__ __launch_bounds__(WF_BLOCK) void k_trace(SceneView S, WavefrontState W, int parity, int pull_slot,
                                                    int32_t* overflow, unsigned long long* counters) {
    __shared__ uint32_t s_stack[LDS_STACK][TRACE_BLOCK];
    __shared__ int s_base;
    const int count = W.counters[C_QCOUNT + parity];
    const int32_t* q = W.queue[parity];
    bool ovf = false;
    TraceCounts c = {0, 0, 0, 0, 0};
    for (;;) {
        int base = pull_chunk(&W.counters[pull_slot], &s_base);
        if (base >= count) break;
        int k = base + (int)threadIdx.x;
        if (k < count) {
            int p = q[k];
            float4 o = W.ray_o[p], d = W.ray_d[p];
            Ray r = make_ray(mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z));
            bool of = false;
            Hit h = trace_closest<COUNT>(S, r, 0.0f, __builtin_inff(), &s_stack[0][threadIdx.x], &of, &c);
            ovf = ovf || of;
            W.hit[p] = make_float4(h.t, bitsf((uint32_t)h.slot), bitsf((uint32_t)h.inst), h.u);
            W.hit_v[p] = h.v;
        }
    }
    if (ovf) atomicOr(overflow, 1);
    if (COUNT) {
        uint32_t v[5] = {c.interior, c.leaf, c.tri, c.entries, c.tlas};
        for (int k = 0; k < 5; ++k) {
            unsigned long long x = v[k];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
            if ((threadIdx.x & 63) == 0) atomicAdd(&counters[k], x);
        }
    }
}

// Material key of a queued path: the hit's material, or `nmat` for a miss.
__device__ __forceinline__ int path_key(const SceneView& S, const WavefrontState& W, int p, int nkeys) {
    float4 h = W.hit[p];
    int slot = (int)fbits(h.y);
    if (slot < 0) return nkeys - 1;
    int m = S.attrs[S.tris[slot].prim].material;
    return m % (nkeys - 1);
}

// Counting sort of the live queue by material key: per-block histograms,
// one exclusive scan (key-major), stable-within-key scatter.
__global__ __launch_bounds__(WF_BLOCK) void k_sort_hist(SceneView S, WavefrontState W, int parity, int nkeys) {
    __shared__ int h[MAX_KEYS];
    const int count = W.counters[C_QCOUNT + parity];
    const int base = blockIdx.x * WF_BLOCK;
    if (base >= count) return;
    for (int k = threadIdx.x; k < nkeys; k += WF_BLOCK) h[k] = 0;
    __syncthreads();
    int i = base + threadIdx.x;
    if (i < count) atomicAdd(&h[path_key(S, W, W.queue[parity][i], nkeys)], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < nkeys; k += WF_BLOCK) W.hist[(size_t)k * gridDim.x + blockIdx.x] = h[k];
}

__global__ __launch_bounds__(1024) void k_sort_scan(WavefrontState W, int parity, int nkeys, int nblocks) {
    // exclusive scan over hist[key][block] (key-major), in place
    __shared__ int s_part[1024];
    const int count = W.counters[C_QCOUNT + parity];
    const int used_blocks = (count + WF_BLOCK - 1) / WF_BLOCK;
    const int n = nkeys * nblocks;
    const int per = (n + 1023) / 1024;
    const int t = threadIdx.x;
    int beg = t * per, end = min(n, beg + per);
    int sum = 0;
    for (int j = beg; j < end; ++j) {
        int b = j % nblocks;
        int v = (b < used_blocks) ? W.hist[j] : 0;
        sum += v;
    }
    s_part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = (t >= off) ? s_part[t - off] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    int run = (t == 0) ? 0 : s_part[t - 1];
    for (int j = beg; j < end; ++j) {
        int b = j % nblocks;
        int v = (b < used_blocks) ? W.hist[j] : 0;
        W.hist[j] = run;
        run += v;
    }
}

__global__ __launch_bounds__(WF_BLOCK) void k_sort_scatter(SceneView S, WavefrontState W, int parity, int nkeys) {
    __shared__ int keys[WF_BLOCK];
    const int count = W.counters[C_QCOUNT + parity];
    const int base = blockIdx.x * WF_BLOCK;
    if (base >= count) return;
    int i = base + threadIdx.x;
    int p = (i < count) ? W.queue[parity][i] : -1;
    int key = (p >= 0) ? path_key(S, W, p, nkeys) : -1;
    keys[threadIdx.x] = key;
    __syncthreads();
    if (p >= 0) {
        int rank = 0;
        for (int j = 0; j < (int)threadIdx.x; ++j) rank += (keys[j] == key) ? 1 : 0;
        int dst = W.hist[(size_t)key * gridDim.x + blockIdx.x] + rank;
        W.sort_tmp[dst] = p;
    }
}

__global__ void k_copy_queue(WavefrontState W, int parity) {
    const int count = W.counters[C_QCOUNT + parity];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) W.queue[parity][i] = W.sort_tmp[i];
}

// One bounce of Li for every live path (integrator.cuh:266-398).
__global__ __launch_bounds__(WF_BLOCK) void k_shade(SceneView S, FrameConst fc, WavefrontState W, int parity,
                                                    int pull_slot) {
    __shared__ int s_base;
    const int count = W.counters[C_QCOUNT + parity];
    const int32_t* q = W.queue[parity];
    int32_t* nq = W.queue[parity ^ 1];
    for (;;) {
        int base = pull_chunk(&W.counters[pull_slot], &s_base);
        if (base >= count) break;
        int k = base + (int)threadIdx.x;
        if (k >= count) continue;
        const int p = q[k];
        float4 o4 = W.ray_o[p], d4 = W.ray_d[p], b4 = W.beta[p], l4 = W.L[p];
        Ray ray = make_ray(mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z));
        uint32_t seed = fbits(d4.w);
        v3 beta = mk3(b4.x, b4.y, b4.z);
        float p_b = b4.w;
        v3 L = mk3(l4.x, l4.y, l4.z);
        uint32_t st = fbits(l4.w);
        int32_t depth = (int32_t)(st & 0xFFFFu);
        bool any_nonspec = (st >> 16) & 1u;
        seed = seed + (uint32_t)depth;
        const bool first = (depth == 0);
        float4 h4 = W.hit[p];
        Hit h;
        h.t = h4.x; h.slot = (int32_t)fbits(h4.y); h.inst = (int32_t)fbits(h4.z); h.u = h4.w; h.v = W.hit_v[p];
        bool alive = false;
        if (h.slot < 0) {
            v3 sky = miss_radiance(fc, ray.d, first);
            if (first) {
                W.vis_a[p] = make_float4(sky.x, sky.y, sky.z, __builtin_inff());
                W.vis_n[p] = make_float4(0.0f, 0.0f, 0.0f, bitsf((uint32_t)-1));
                W.vis_p[p] = make_float4(0.0f, 0.0f, 0.0f, bitsf((uint32_t)-1));
            }
            L = L + sky * beta;
        } else {
            Surf s = make_surf(S, h, ray);
            float4 pp = W.prev_pos[p];
            v3 Le, add;
            if (emitted(S, fc, s, first, p_b, mk3(pp.x, pp.y, pp.z), beta, &Le, &add)) L = L + add;
            BSDF b = get_bsdf(S, s);
            if (first) {
                v3 alb = !is_zero(Le) ? Le : b.albedo;
                W.vis_a[p] = make_float4(alb.x, alb.y, alb.z, s.distance);
                W.vis_n[p] = make_float4(s.n.x, s.n.y, s.n.z, bitsf((uint32_t)s.inst));
                W.vis_p[p] = make_float4(s.pos.x, s.pos.y, s.pos.z, bitsf((uint32_t)s.prim));
                b.albedo = mk3(1.0f, 1.0f, 1.0f);
            }
            if (any_nonspec) bsdf_regularize(b);
            if (!(b.alpha < 0.001f)) {
                NeeResult ne = next_event(S, fc, s, b, ray.d, beta, seed);
                if (ne.trace) {
                    int sidx = atomicAdd(&W.counters[C_SHADOW], 1);
                    W.shadow_queue[sidx] = p;
                    W.shadow_o[sidx] = make_float4(ne.org.x, ne.org.y, ne.org.z, ne.tmax);
                    W.shadow_d[sidx] = make_float4(ne.dir.x, ne.dir.y, ne.dir.z, 0.0f);
                    W.shadow_c[sidx] = make_float4(ne.unoccl.x, ne.unoccl.y, ne.unoccl.z, 0.0f);
                    W.shadow_z[sidx] = make_float4(ne.occl.x, ne.occl.y, ne.occl.z, 0.0f);
                } else {
                    L = L + ne.occl;
                }
            }
            const bool at_max = (depth == fc.max_depth);
            depth++;
            if (!at_max) {
                v3 wo = -ray.d;
                v2 u2, X2;
                u2.x = rng1(seed); u2.y = rng1(seed);
                X2.x = rng1(seed); X2.y = rng1(seed);
                BSample bs = bsdf_sample(b, wo, u2, X2);
                if (!is_zero(bs.f)) {
                    beta = beta * (bs.f * absdot(s.n, bs.wi) / bs.pdf);
                    p_b = bsdf_pdf(b, wo, bs.wi);
                    any_nonspec = any_nonspec || !((bs.scatter & SC_SPECULAR) != 0);
                    v3 org = (bs.scatter & SC_TRANSMITTED) ? s.pos - (s.n * KPT_HIT_EPS) : s.pos + (s.n * KPT_HIT_EPS);
                    ray = make_ray(org, bs.wi);
                    alive = true;
                    v3 rr = beta * 1.0f;
                    if (max_comp(rr) < 1.0f && depth > 1) {
                        float qq = fmaxf_(0.0f, 1.0f - max_comp(rr));
                        if (rng1(seed) < qq) alive = false;
                        else beta = beta / (1.0f - qq);
                    }
                    W.prev_pos[p] = make_float4(s.pos.x, s.pos.y, s.pos.z, 0.0f);
                    if (alive && is_zero(beta)) alive = false;
                }
            }
        }
        W.L[p] = make_float4(L.x, L.y, L.z, bitsf((uint32_t)depth | (any_nonspec ? 0x10000u : 0u)));
        W.ray_d[p] = make_float4(ray.d.x, ray.d.y, ray.d.z, bitsf(seed));
        if (alive) {
            W.ray_o[p] = make_float4(ray.o.x, ray.o.y, ray.o.z, o4.w);
            W.beta[p] = make_float4(beta.x, beta.y, beta.z, p_b);
            int ni = atomicAdd(&W.counters[C_QCOUNT + (parity ^ 1)], 1);
            nq[ni] = p;
        }
    }
}

// Visibility of the deferred next-event estimates; adds their term to L.
__global__ __launch_bounds__(WF_BLOCK) void k_shadow(SceneView S, WavefrontState W, int pull_slot, int32_t* overflow) {
    __shared__ uint32_t s_stack[LDS_STACK][TRACE_BLOCK];
    __shared__ int s_base;
    const int count = W.counters[C_SHADOW];
    bool ovf = false;
    for (;;) {
        int base = pull_chunk(&W.counters[pull_slot], &s_base);
        if (base >= count) break;
        int k = base + (int)threadIdx.x;
        if (k < count) {
            float4 o = W.shadow_o[k], d = W.shadow_d[k];
            Ray r = make_ray(mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z));
            bool of = false;
            bool occluded = trace_any<false>(S, r, 0.0f, o.w, &s_stack[0][threadIdx.x], &of, nullptr);
            ovf = ovf || of;
            int p = W.shadow_queue[k];
            float4 c = occluded ? W.shadow_z[k] : W.shadow_c[k];
            float4 l = W.L[p];
            W.L[p] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w);
        }
    }
    if (ovf) atomicOr(overflow, 1);
}

// clampOutput, G/V-buffer, static accumulation (kernels.cu:411-450).
__global__ __launch_bounds__(WF_BLOCK) void k_finish(SceneView S, FrameConst fc, int32_t frame, WavefrontState W,
                                                     FrameBuffers fb) {
    const int64_t npx = (int64_t)fc.width * fc.height;
    const int64_t i = (int64_t)blockIdx.x * WF_BLOCK + threadIdx.x;
    if (i >= npx) return;
    float4 l = W.L[i];
    v3 L = (mk3(1.0f, 1.0f, 1.0f) * 1.0f) * mk3(l.x, l.y, l.z);
    L = clamp_output(L);
    float4 va = W.vis_a[i], vn = W.vis_n[i], vp = W.vis_p[i];
    int inst = (int)fbits(vn.w), prim = (int)fbits(vp.w);
    fb.gbuffer[i] = pack_gbuffer(mk3(va.x, va.y, va.z), mk3(vn.x, vn.y, vn.z), va.w);
    v2 vel = velocity(S, fc, inst, mk3(vp.x, vp.y, vp.z));
    fb.vbuffer[i] = make_float4(vel.x, vel.y, (float)prim, (float)inst);
    const float4 a = fb.accum[i];
    v3 acc = mk3(a.x, a.y, a.z) + L;
    fb.accum[i] = make_float4(acc.x, acc.y, acc.z, 1.0f);
    v3 est = acc / (float)(frame + 1);
    fb.render[i] = make_float4(est.x, est.y, est.z, 1.0f);
}

int g_num_cus = 0;

}  // namespace

hipError_t wavefront_frame(hipStream_t s, const SceneView& S, const FrameConst& fc, int32_t frame,
                           const FrameBuffers& fb, WavefrontState& W, int32_t* overflow, int32_t num_materials,
                           int32_t sort_by_material, WfInstrument* ins) {
    if (ins) ins->launches = 0;
    if (g_num_cus == 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) g_num_cus = prop.multiProcessorCount;
        if (g_num_cus <= 0) g_num_cus = 256;
    }
    const int64_t npx = (int64_t)fc.width * fc.height;
    const unsigned gblocks = (unsigned)((npx + WF_BLOCK - 1) / WF_BLOCK);
    // persistent grids: 8 workgroups of 256 lanes per CU
    const unsigned pgrid = (unsigned)std::min<int64_t>((int64_t)g_num_cus * 8, gblocks);
    const int nkeys = std::min(MAX_KEYS, num_materials + 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, W.counters, 0, C_PULL + 3 * MAX_BOUNCE_SLOTS);
    hipLaunchKernelGGL(k_generate, dim3(gblocks), dim3(WF_BLOCK), 0, s, fc, frame, W);
    const int max_bounces = std::min(fc.max_depth + 1, MAX_BOUNCE_SLOTS);
    for (int bounce = 0; bounce < max_bounces; ++bounce) {
        const int parity = bounce & 1;
        // next queue and shadow count start at zero
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, W.counters + C_QCOUNT + (parity ^ 1), 0, 1);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, W.counters + C_SHADOW, 0, 1);
        const bool timed = ins && ins->launches < WF_MAX_TRACE_LAUNCHES;
        if (timed) (void)hipEventRecord(ins->begin[ins->launches], s);
        if (ins && ins->counters) {
            hipLaunchKernelGGL(k_trace<true>, dim3(pgrid), dim3(WF_BLOCK), 0, s, S, W, parity, C_PULL + 3 * bounce,
                               overflow, ins->counters);
        } else {
            hipLaunchKernelGGL(k_trace<false>, dim3(pgrid), dim3(WF_BLOCK), 0, s, S, W, parity, C_PULL + 3 * bounce,
                               overflow, (unsigned long long*)nullptr);
        }
        if (timed) (void)hipEventRecord(ins->end[ins->launches++], s);
        if (sort_by_material && nkeys > 1) {
            hipLaunchKernelGGL(k_sort_hist, dim3(gblocks), dim3(WF_BLOCK), 0, s, S, W, parity, nkeys);
            hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, s, W, parity, nkeys, (int)gblocks);
            hipLaunchKernelGGL(k_sort_scatter, dim3(gblocks), dim3(WF_BLOCK), 0, s, S, W, parity, nkeys);
            hipLaunchKernelGGL(k_copy_queue, dim3(gblocks), dim3(WF_BLOCK), 0, s, W, parity);
        }
        hipLaunchKernelGGL(k_shade, dim3(pgrid), dim3(WF_BLOCK), 0, s, S, fc, W, parity, C_PULL + 3 * bounce + 1);
        hipLaunchKernelGGL(k_shadow, dim3(pgrid), dim3(WF_BLOCK), 0, s, S, W, C_PULL + 3 * bounce + 2, overflow);
    }
    hipLaunchKernelGGL(k_finish, dim3(gblocks), dim3(WF_BLOCK), 0, s, S, fc, frame, W, fb);
    return hipGetLastError();
}

}  // namespace kpt
