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
#include <rocprim/device/device_radix_sort.hpp>

#include "kpt_integrator.h"
#include "kpt_launch.h"
#include "kpt_temporal.h"

namespace kpt {

namespace {

constexpr int WF_BLOCK = 256;
// Occupancy targets in waves per SIMD (512 VGPRs per lane per SIMD): the
// traversal kernels at 5 (at most 96 VGPRs; they need 100 unconstrained and
// fit without spilling), the path-per-lane tail at 4 (128 VGPRs, from 166).
// Both measured faster than the unconstrained allocation on the instanced
// 1080p workload.
#ifndef KPT_WPE_TRACE
#define KPT_WPE_TRACE __attribute__((amdgpu_waves_per_eu(5)))
#endif
#ifndef KPT_WPE_TAIL
#define KPT_WPE_TAIL __attribute__((amdgpu_waves_per_eu(4)))
#endif
#ifndef KPT_WPE_SHADE
#define KPT_WPE_SHADE
#endif
constexpr int MAX_KEYS = 64;
// counters[] slots (zeroed once per frame by a fill kernel)
constexpr int C_QCOUNT = 0;      // + bounce parity: [0],[1] live-queue sizes
constexpr int C_SHADOW = 2;      // shadow-queue size of the current bounce
constexpr int C_PULL = 4;        // 4..(4+3*MAX_BOUNCE): pull counters per stage and bounce
constexpr int MAX_BOUNCE_SLOTS = 40;
static_assert(C_PULL + 3 * MAX_BOUNCE_SLOTS <= WF_COUNTER_SLOTS, "counter slots");
// A slot's word in W.counters: slots lie WF_COUNTER_STRIDE words apart, so
// the queue claims and appends of one kernel (up to three counters, each
// taking an atomic per wave and chunk) go to different cache lines and L2
// channels instead of queueing behind one another on one line.
KD int cs(int slot) { return slot * WF_COUNTER_STRIDE; }

KD uint32_t fbits(float f) { return __builtin_bit_cast(uint32_t, f); }
KD float bitsf(uint32_t u) { return __builtin_bit_cast(float, u); }

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

// One atomic per wave: claims `n` consecutive queue entries (wave-uniform result).
__device__ __forceinline__ int wave_claim(int32_t* ctr, int n) {
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(ctr, n);
    return __shfl(base, 0, WAVE);
}

// Wave-local work source over a global queue of `count` entries: a wave claims
// `chunk` entries at a time and hands them to its idle lanes (once at least
// `refill_min` are idle), so a lane whose ray finishes starts another without
// waiting for the rest of the wave.
struct WaveFeeder {
    int pos, end;        // wave-uniform claimed range [pos, end)
    bool drained;
    int chunk, refill_min;
    __device__ void init(const TraceTuning& t) {
        pos = 0; end = 0; drained = false; chunk = t.chunk; refill_min = t.refill_min;
    }
    // Returns the queue position for this lane (when `idle`) or -1.
    __device__ int feed(int32_t* ctr, int count, bool idle) {
        const uint64_t m = __ballot(idle);
        const int n = __popcll(m);
        if (n == 0 || (n < refill_min && n < WAVE)) return -1;
        if (pos >= end && !drained) {
            int b = wave_claim(ctr, chunk);
            if (b >= count) drained = true;
            else { pos = b; end = min(b + chunk, count); }
        }
        if (pos >= end) return -1;
        const int k = pos + __popcll(m & lanes_below());
        pos = min(pos + n, end);
        return (idle && k < end) ? k : -1;
    }
    __device__ bool exhausted() const { return drained && pos >= end; }
};

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
    // queue order: 8x8 pixel tiles, one per wave, when the frame divides into
    // them (a wave's camera rays then share more nodes than a 64-pixel row);
    // paths stay indexed by pixel, only the processing order changes
    int64_t slot = i;
    if ((fc.width & 7) == 0 && (fc.height & 7) == 0) {
        const int64_t tile = (int64_t)(py >> 3) * (fc.width >> 3) + (px >> 3);
        slot = tile * 64 + (py & 7) * 8 + (px & 7);
    }
    W.queue[0][slot] = (int32_t)i;
    if (i == 0) W.counters[cs(C_QCOUNT)] = (int32_t)npx;
}

// Multi-view generation: path i belongs to the view whose pixel range holds
// it (binary search over the views' offsets); the seed and the camera ray are
// those of that view's own pixel, so a view renders exactly as it would alone.
__global__ __launch_bounds__(WF_BLOCK) void k_generate_views(FrameConst fc, int32_t frame, WavefrontState W,
                                                             ViewBatch vb) {
    const int64_t i = (int64_t)blockIdx.x * WF_BLOCK + threadIdx.x;
    if (i >= vb.total) return;
    int lo = 0, hi = vb.nviews - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (vb.views[mid].offset <= i) lo = mid; else hi = mid - 1;
    }
    const ViewConst& v = vb.views[lo];
    const int64_t j = i - v.offset;
    const int px = (int)(j % v.width), py = (int)(j / v.width);
    uint32_t seed = init_seed(px, py, v.width, frame, fc.seed_mode);
    (void)rng1(seed);
    (void)rng1(seed);
    Ray r = camera_ray_m(v.inv_view, v.inv_proj, v.cam_pos, v.width, v.height, px, py);
    W.ray_o[i] = make_float4(r.o.x, r.o.y, r.o.z, bitsf((uint32_t)i));
    W.ray_d[i] = make_float4(r.d.x, r.d.y, r.d.z, bitsf(seed));
    W.beta[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    W.L[i] = make_float4(0.0f, 0.0f, 0.0f, bitsf(0u));
    W.prev_pos[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    W.queue[0][i] = (int32_t)i;
    if (i == 0) W.counters[cs(C_QCOUNT)] = (int32_t)vb.total;
}

// Multi-view finish: clampOutput and the running sum of the samples.
__global__ __launch_bounds__(WF_BLOCK) void k_finish_views(WavefrontState W, ViewBatch vb) {
    const int64_t i = (int64_t)blockIdx.x * WF_BLOCK + threadIdx.x;
    if (i >= vb.total) return;
    const float4 l = W.L[i];
    v3 L = (mk3(1.0f, 1.0f, 1.0f) * 1.0f) * mk3(l.x, l.y, l.z);
    L = clamp_output(L);
    const float4 a = vb.out[i];
    v3 acc = mk3(a.x, a.y, a.z) + L;
    vb.out[i] = make_float4(acc.x, acc.y, acc.z, 1.0f);
}

__global__ __launch_bounds__(WF_BLOCK) void k_views_mean(float4* out, int64_t n, int32_t spp) {
    const int64_t i = (int64_t)blockIdx.x * WF_BLOCK + threadIdx.x;
    if (i >= n) return;
    const float4 a = out[i];
    v3 est = mk3(a.x, a.y, a.z) / (float)spp;
    out[i] = make_float4(est.x, est.y, est.z, 1.0f);
}

// DebugData heat map colour, kernels.cu:442-443: green 0.02 per BLAS node
// pop, blue 0.05 per TLAS node pop.
__global__ __launch_bounds__(WF_BLOCK) void k_heat_colour(const int32_t* heat, int64_t n, float4* rgba) {
    const int64_t i = (int64_t)blockIdx.x * WF_BLOCK + threadIdx.x;
    if (i >= n) return;
    const float b = (float)heat[2 * i], t = (float)heat[2 * i + 1];
    rgba[i] = make_float4(0.0f, b * 0.02f, t * 0.05f, 1.0f);
}

// ---- spatial ray sort: rays with nearby origins (and, for continuation rays,
// the same direction octant) end up in the same wave, so a wave's traversals
// share nodes and cache lines.  Results do not depend on the order.
constexpr uint32_t RKEY_INVALID = 1u << 15;   // sorts after every valid 15-bit key
constexpr int RKEY_BITS = 16;

KD uint32_t spread4(uint32_t v) {   // bits b3..b0 -> b3 _ _ b2 _ _ b1 _ _ b0
    return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4) | ((v & 8u) << 6);
}
KD uint32_t ray_key(const FrameConst& fc, float4 o, float4 d, bool with_dir) {
    auto cell = [&](float x, int a) {
        float t = (x - fc.scene_lo[a]) * fc.scene_inv[a];
        t = fminf_(fmaxf_(t, 0.0f), 0.9999f);
        return (uint32_t)(t * 16.0f);
    };
    uint32_t m = spread4(cell(o.x, 0)) | (spread4(cell(o.y, 1)) << 1) | (spread4(cell(o.z, 2)) << 2);
    if (with_dir) m |= ((d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u)) << 12;
    return m;
}

__global__ __launch_bounds__(WF_BLOCK) void k_ray_keys(FrameConst fc, WavefrontState W, int parity, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * WF_BLOCK + threadIdx.x;
    if (i >= n) return;
    const int count = W.counters[cs(C_QCOUNT + parity)];
    if (i < count) {
        const int p = W.queue[parity][i];
        W.rkeys[0][i] = ray_key(fc, W.ray_o[p], W.ray_d[p], true);
        W.rvals[0][i] = p;
    } else {
        W.rkeys[0][i] = RKEY_INVALID;
        W.rvals[0][i] = -1;
    }
}

__global__ __launch_bounds__(WF_BLOCK) void k_shadow_keys(FrameConst fc, WavefrontState W, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * WF_BLOCK + threadIdx.x;
    if (i >= n) return;
    const int count = W.counters[cs(C_SHADOW)];
    if (i < count) {
        W.rkeys[0][i] = ray_key(fc, W.shadow_o[i], W.shadow_d[i], true);
        W.rvals[0][i] = (int32_t)i;
    } else {
        W.rkeys[0][i] = RKEY_INVALID;
        W.rvals[0][i] = -1;
    }
}

// Counting mode: one closest-hit trace's node pops added to its pixel's
// DebugData, blas_hits and tlas_hits (blas.cuh:140, tlas.cu:87).
__device__ __forceinline__ void heat_add(int32_t* heat, float pixel_bits, uint32_t blas_pops, uint32_t tlas_pops) {
    int32_t* h = heat + 2 * (int64_t)fbits(pixel_bits);
    atomicAdd(&h[0], (int32_t)blas_pops);
    atomicAdd(&h[1], (int32_t)tlas_pops);
}

// Closest hit of every live path.  COUNT accumulates traversal events (and,
// when W.heat is set, each pixel's node pops).
template <bool COUNT>
__global__ __launch_bounds__(WF_BLOCK) KPT_WPE_TRACE void k_trace(SceneView S, WavefrontState W, int parity, int pull_slot,
                                                    int32_t* overflow, unsigned long long* counters,
                                                    TraceTuning tune, const int32_t* sorted) {
    extern __shared__ uint32_t s_dyn[];   // tune.stack x blockDim.x words
    uint32_t spill[SPILL_STACK];
    const int count = W.counters[cs(C_QCOUNT + parity)];
    const int32_t* q = sorted ? sorted : W.queue[parity];
    int32_t* ctr = &W.counters[cs(pull_slot)];
    const StackShape lds{s_dyn + threadIdx.x, (int)blockDim.x, tune.stack};
    bool ovf = false;
    TraceCounts c = {0, 0, 0, 0, 0};
    if (!tune.refill) {
        // one ray per lane to completion; a wave claims tune.chunk entries
        // (a multiple of 64) per atomic and runs them 64 at a time
        const int chunk = (tune.chunk / WAVE) * WAVE;
        for (;;) {
            const int base = wave_claim(ctr, chunk);
            if (base >= count) break;
            for (int off = 0; off < chunk && base + off < count; off += WAVE) {
                const int k = base + off + lane_id();
                if (k >= count) continue;
                const int p = q[k];
                const float4 o = W.ray_o[p], d = W.ray_d[p];
                const uint32_t b0 = c.interior + c.leaf, t0 = c.tlas;
                ClosestTrav tr;
                if (tr.init(S, make_ray(mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z)), 0.0f, __builtin_inff(), lds,
                            spill)) {
                    while (tr.template step<COUNT>(S, &c)) {}
                }
                if (COUNT && W.heat) heat_add(W.heat, o.w, c.interior + c.leaf - b0, c.tlas - t0);
                ovf = ovf || tr.st.overflow;
                W.hit[p] = make_float4(tr.h.t, bitsf((uint32_t)tr.h.slot), bitsf((uint32_t)tr.h.inst), tr.h.u);
                W.hit_v[p] = tr.h.v;
            }
        }
    } else {
        WaveFeeder feeder;
        feeder.init(tune);
        ClosestTrav tr;
        int p = -1;
        bool busy = false;
        uint32_t b0 = 0u, t0 = 0u;   // counting mode: the lane's counts when its ray started
        float pix = 0.0f;
        for (;;) {
            const int k = feeder.feed(ctr, count, !busy);
            bool done_now = false;
            if (k >= 0) {
                p = q[k];
                const float4 o = W.ray_o[p], d = W.ray_d[p];
                if (COUNT) { b0 = c.interior + c.leaf; t0 = c.tlas; pix = o.w; }
                busy = tr.init(S, make_ray(mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z)), 0.0f, __builtin_inff(), lds,
                               spill);
                done_now = !busy;   // empty scene or root box missed
            }
            if (!__any(busy || done_now)) {   // wave-uniform
                if (feeder.exhausted()) break;
                continue;
            }
            if (busy) {
                busy = tr.template step<COUNT>(S, &c);
                done_now = !busy;
            }
            if (done_now) {
                ovf = ovf || tr.st.overflow;
                if (COUNT && W.heat) heat_add(W.heat, pix, c.interior + c.leaf - b0, c.tlas - t0);
                W.hit[p] = make_float4(tr.h.t, bitsf((uint32_t)tr.h.slot), bitsf((uint32_t)tr.h.inst), tr.h.u);
                W.hit_v[p] = tr.h.v;
            }
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

// Counting sort of the live queue by material key (misses form their own
// key), so a shading wave runs one material's code: block-aggregated key
// totals, an exclusive scan of the (few) keys, then a scatter in which each
// block reserves its range per key with one atomic.  Order inside a key is
// free: paths are independent.
// W.hist layout: [0, MAX_KEYS) totals, [MAX_KEYS, 2*MAX_KEYS) running cursors.
__global__ __launch_bounds__(WF_BLOCK) void k_sort_hist(SceneView S, WavefrontState W, int parity, int nkeys) {
    __shared__ int h[MAX_KEYS];
    const int count = W.counters[cs(C_QCOUNT + parity)];
    for (int k = threadIdx.x; k < nkeys; k += WF_BLOCK) h[k] = 0;
    __syncthreads();
    for (int i = blockIdx.x * WF_BLOCK + threadIdx.x; i < count; i += gridDim.x * WF_BLOCK)
        atomicAdd(&h[path_key(S, W, W.queue[parity][i], nkeys)], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < nkeys; k += WF_BLOCK)
        if (h[k]) atomicAdd(&W.hist[k], h[k]);
}

__global__ void k_sort_offsets(WavefrontState W, int nkeys) {
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = 0; k < nkeys; ++k) { W.hist[MAX_KEYS + k] = run; run += W.hist[k]; }
    }
}

__global__ __launch_bounds__(WF_BLOCK) void k_sort_scatter(SceneView S, WavefrontState W, int parity, int nkeys) {
    __shared__ int h[MAX_KEYS];
    __shared__ int base[MAX_KEYS];
    const int count = W.counters[cs(C_QCOUNT + parity)];
    for (int b0 = blockIdx.x * WF_BLOCK; b0 < count; b0 += gridDim.x * WF_BLOCK) {
        for (int k = threadIdx.x; k < nkeys; k += WF_BLOCK) h[k] = 0;
        __syncthreads();
        const int i = b0 + threadIdx.x;
        const int p = (i < count) ? W.queue[parity][i] : -1;
        const int key = (p >= 0) ? path_key(S, W, p, nkeys) : 0;
        const int rank = (p >= 0) ? atomicAdd(&h[key], 1) : 0;
        __syncthreads();
        for (int k = threadIdx.x; k < nkeys; k += WF_BLOCK)
            base[k] = h[k] ? atomicAdd(&W.hist[MAX_KEYS + k], h[k]) : 0;
        __syncthreads();
        if (p >= 0) W.sort_tmp[base[key] + rank] = p;
        __syncthreads();
    }
}

// One bounce of Li for every live path (integrator.cuh:266-398), one queue
// entry per lane.  A workgroup reserves its shadow rays and its surviving
// paths with one atomic per queue after its lanes have shaded (a prefix over
// its waves' ballots in LDS): the appends of every workgroup target the same
// two counters, and one atomic per 256 lanes instead of one per wave and
// queue keeps that address's queue in L2 short.
__global__ __launch_bounds__(WF_BLOCK) KPT_WPE_SHADE void k_shade(SceneView S, FrameConst fc, WavefrontState W, int parity,
                                                    const int32_t* q, CacheBuffers cb, CacheFrame cf) {
    constexpr int NW = WF_BLOCK / WAVE;
    __shared__ int s_off[2][NW + 1];   // per queue: the waves' offsets, then the workgroup's base
    const int count = W.counters[cs(C_QCOUNT + parity)];
    const int first_k = (int)blockIdx.x * WF_BLOCK;
    if (first_k >= count) return;   // workgroup-uniform
    const int k = first_k + (int)threadIdx.x;
    bool want_shadow = false, alive = false;
    int p = -1;
    float4 sh_o, sh_d, sh_c, sh_z;
    if (k < count) {
        p = q[k];
        float4 o4 = W.ray_o[p], d4 = W.ray_d[p], b4 = W.beta[p], l4 = W.L[p];
        Ray ray = make_ray(mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z));
        uint32_t seed = fbits(d4.w);
        v3 beta = mk3(b4.x, b4.y, b4.z);
        float p_b = b4.w;
        v3 L = mk3(l4.x, l4.y, l4.z);
        uint32_t st = fbits(l4.w);
        int32_t depth = (int32_t)(st & 0xFFFFu);
        bool any_nonspec = (st >> 16) & 1u;
        bool wide_in = (st >> 17) & 1u;      // this vertex was reached through a wide lobe
        bool cached = (st >> 18) & 1u;       // the path has met its cache vertex (training record made)
        seed = seed + (uint32_t)depth;
        const bool first = (depth == 0);
        float4 h4 = W.hit[p];
        Hit h;
        h.t = h4.x; h.slot = (int32_t)fbits(h4.y); h.inst = (int32_t)fbits(h4.z); h.u = h4.w; h.v = W.hit_v[p];
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
            // radiance cache at the path's cache vertex (cache_vertex)
            const CacheVertex cv = cache_vertex(fc, cb, cf, p, depth, wide_in, cached, s, b, ray.d, beta);
            const bool to_cache = cv.to_cache;
            int32_t trec = cv.trec;
            const float sun_p = cv.sun_p;
            if (trec >= 0 && !cf.nee) {
                cb.t_lbefore[trec] = make_float4(L.x, L.y, L.z, 0.0f);
                trec = -1;
            }
            if (!(to_cache && !cf.nee) && !(b.alpha < 0.001f)) {
                NeeResult ne = next_event(S, fc, s, b, ray.d, beta, seed, sun_p);
                if (ne.trace) {
                    want_shadow = true;
                    sh_o = make_float4(ne.org.x, ne.org.y, ne.org.z, ne.tmax);
                    sh_d = make_float4(ne.dir.x, ne.dir.y, ne.dir.z, bitsf((uint32_t)trec));
                    sh_c = make_float4(ne.unoccl.x, ne.unoccl.y, ne.unoccl.z, 0.0f);
                    sh_z = make_float4(ne.occl.x, ne.occl.y, ne.occl.z, 0.0f);
                } else {
                    L = L + ne.occl;
                }
            }
            // a deferred estimate rewrites this once its visibility is known
            if (trec >= 0) cb.t_lbefore[trec] = make_float4(L.x, L.y, L.z, 0.0f);
            const bool at_max = (depth == fc.max_depth);
            depth++;
            if (!at_max && !to_cache) {
                v3 wo = -ray.d;
                v2 u2, X2;
                u2.x = rng1(seed); u2.y = rng1(seed);
                X2.x = rng1(seed); X2.y = rng1(seed);
                BSample bs = bsdf_sample(b, wo, u2, X2);
                if (!is_zero(bs.f)) {
                    beta = beta * (bs.f * absdot(s.n, bs.wi) / bs.pdf);
                    p_b = bsdf_pdf(b, wo, bs.wi);
                    any_nonspec = any_nonspec || !((bs.scatter & SC_SPECULAR) != 0);
                    wide_in = ((bs.scatter & SC_DIFFUSE) != 0) || b.roughness >= cf.wide_roughness;
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
        W.L[p] = make_float4(L.x, L.y, L.z, bitsf((uint32_t)depth | (any_nonspec ? 0x10000u : 0u) |
                                                  (wide_in ? 0x20000u : 0u) | (cached ? 0x40000u : 0u)));
        W.ray_d[p] = make_float4(ray.d.x, ray.d.y, ray.d.z, bitsf(seed));
        if (alive) {
            W.ray_o[p] = make_float4(ray.o.x, ray.o.y, ray.o.z, o4.w);
            W.beta[p] = make_float4(beta.x, beta.y, beta.z, p_b);
        }
    }
    const uint64_t m_sh = __ballot(want_shadow), m_al = __ballot(alive);
    const int wave = (int)(threadIdx.x / WAVE);
    if (lane_id() == 0) {
        s_off[0][wave] = __popcll(m_sh);
        s_off[1][wave] = __popcll(m_al);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int j = (int)threadIdx.x;
        int run = 0;
        for (int w = 0; w < NW; ++w) {
            const int c = s_off[j][w];
            s_off[j][w] = run;
            run += c;
        }
        s_off[j][NW] = run ? atomicAdd(&W.counters[cs(j == 0 ? C_SHADOW : C_QCOUNT + (parity ^ 1))], run) : 0;
    }
    __syncthreads();
    if (want_shadow) {
        const int sidx = s_off[0][NW] + s_off[0][wave] + __popcll(m_sh & lanes_below());
        W.shadow_queue[sidx] = p;
        W.shadow_o[sidx] = sh_o;
        W.shadow_d[sidx] = sh_d;
        W.shadow_c[sidx] = sh_c;
        W.shadow_z[sidx] = sh_z;
    }
    if (alive) W.queue[parity ^ 1][s_off[1][NW] + s_off[1][wave] + __popcll(m_al & lanes_below())] = p;
}

// The rest of every path still alive, one path per lane (li_bounce from the
// path's state in HBM): after the first bounces few paths remain, each needing
// several dependent trace/shade/shadow rounds, and per-bounce launches would
// be bound by the latency of their few rays.  Same arithmetic and random
// numbers as the stage kernels, so the radiance is unchanged.  Lanes refill
// at vertex boundaries: a lane whose path ends takes the next queued path
// before the wave's next vertex, so a wave is not held by its longest path.
// A path whose ray escapes leaves its sky term to k_sky: the lane appends
// (path, direction, throughput) to the sky queue (the shadow-ray arrays,
// unused while the tail runs; count in C_SHADOW, cleared before the launch)
// instead of evaluating the sky with the few lanes that missed.
// Parked traces (suspend > 0): a vertex's closest-hit trace holds the whole
// wave until its longest ray ends, while most lanes have finished.  Once
// fewer than `suspend` lanes are still tracing (and the wave can still claim
// new paths), the wave stops the loop: each unfinished trace is parked — its
// path state written back to the path arrays as a queued path's, its
// traversal state and stack to one of the wave's slots in W.tpool — and its
// lane goes on as an idle lane.  Idle lanes take parked traces before new
// paths, load the path as a claimed one and continue the traversal from the
// state it stopped in, so every ray runs the same steps in the same order and
// the image is unchanged.  Each parked trace has run at least one step since
// it was begun or taken back, and a wave that has drained the queue parks
// nothing, so every trace ends.
// CACHE: the radiance cache's hooks at each path's cache vertex (li_shade).
constexpr int TAIL_POOL = 64;         // parked traces per wave
constexpr int TAIL_SLOT_HEAD = 4;     // float4 of a parked trace's state before its stack
constexpr int TAIL_SAVE_STACK = 48;   // stack entries a parked trace may hold
constexpr int TAIL_SLOT_F4 = TAIL_SLOT_HEAD + TAIL_SAVE_STACK / 4;   // float4 per parked trace
static_assert(TAIL_POOL * TAIL_SLOT_F4 * 16 == TAIL_POOL_BYTES, "pool layout");
constexpr int TAIL_CHECK = 16;        // trace steps per lane between parking checks

template <bool COUNT, bool CACHE>
__global__ __launch_bounds__(WF_BLOCK) KPT_WPE_TAIL void k_tail(SceneView S, FrameConst fc, WavefrontState W, int parity,
                                                   int pull_slot, int32_t* overflow,
                                                   unsigned long long* counters, const int32_t* order,
                                                   CacheBuffers cb, CacheFrame cf, int stack_depth, int suspend) {
    extern __shared__ uint32_t s_dyn[];   // stack_depth x blockDim.x words
    const StackShape lds{s_dyn + threadIdx.x, (int)blockDim.x, stack_depth};
    const int count = W.counters[cs(C_QCOUNT + parity)];
    const int32_t* q = order ? order : W.queue[parity];
    int32_t* ctr = &W.counters[cs(pull_slot)];
    const int64_t gwave = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    float4* pool = (suspend > 0 && W.tpool && gwave < W.tpool_waves)
                       ? W.tpool + gwave * (int64_t)(TAIL_POOL * TAIL_SLOT_F4) : nullptr;
    int npool = 0;   // wave-uniform: parked traces in the wave's slots [0, npool)
    bool ovf = false;
    bool active = false, drained = false;   // drained: wave-uniform
    int p = -1;
    uint32_t st = 0u, seed = 0u;
    PathState ps;
    VisOut vs;   // written only at depth 0, which never reaches this kernel
    TraceCounts c = {0, 0, 0, 0, 0};
    uint32_t n_paths = 0, n_vertices = 0;
    int32_t* heat_px = nullptr;   // counting mode: the lane's pixel in W.heat
    uint32_t spill[SPILL_STACK];
    for (;;) {
        uint64_t idle = __ballot(!active);
        bool fresh = false, resumed = false;
        int slot = 0;
        // wave-uniform: parked traces first, the newest first
        if (idle != 0ull && npool > 0) {
            const int r = min(npool, __popcll(idle));
            const int rank = __popcll(idle & lanes_below());
            if (!active && rank < r) {
                slot = npool - 1 - rank;
                p = (int)fbits(pool[slot * TAIL_SLOT_F4].x);
                resumed = true;
            }
            npool -= r;
            idle = __ballot(!active && !resumed);
        }
        // wave-uniform: claim one queue entry per idle lane
        if (idle != 0ull && !drained) {
            const int base = wave_claim(ctr, __popcll(idle));
            if (base + __popcll(idle) >= count) drained = true;
            const int k = base + __popcll(idle & lanes_below());
            if (!active && !resumed && k < count) {
                p = q[k];
                fresh = true;
                if (COUNT) ++n_paths;
            }
        }
        if (fresh || resumed) {
            const float4 o4 = W.ray_o[p], d4 = W.ray_d[p], b4 = W.beta[p], l4 = W.L[p], pp = W.prev_pos[p];
            ps.ray = make_ray(mk3(o4.x, o4.y, o4.z), mk3(d4.x, d4.y, d4.z));
            ps.L = mk3(l4.x, l4.y, l4.z);
            ps.beta = mk3(b4.x, b4.y, b4.z);
            ps.p_b = b4.w;
            st = fbits(l4.w);
            ps.depth = (int32_t)(st & 0xFFFFu);
            ps.any_nonspec = (st >> 16) & 1u;
            ps.wide_in = (st >> 17) & 1u;
            ps.cached = (st >> 18) & 1u;
            ps.prev_pos = mk3(pp.x, pp.y, pp.z);
            seed = fbits(d4.w);
            active = true;
            if (COUNT && W.heat) heat_px = W.heat + 2 * (int64_t)fbits(o4.w);
        }
        if (!__any(active)) break;   // wave-uniform: queue drained, nothing parked, every lane idle
        // the vertex's closest-hit trace: begun here, or taken back where it was parked
        bool missed = false, parked = false;
        if (active) {
            bool o = false, more = false;
            if (resumed || !is_zero(ps.beta)) {
                ClosestTrav tr;
                bool busy = true;
                if (resumed) {
                    // the world ray is the path's ray (as the trace was begun with it), the
                    // object ray recomputed from it as the instance step computed it
                    const float4* sl = pool + slot * TAIL_SLOT_F4;
                    const float4 a0 = sl[0], a1 = sl[1], a2 = sl[2], a3 = sl[3];
                    tr.w = ps.ray;
                    tr.tmin = a0.y; tr.tmax = a0.z; tr.h.t = a0.w;
                    tr.h.slot = (int32_t)fbits(a1.x); tr.h.inst = (int32_t)fbits(a1.y); tr.h.u = a1.z; tr.h.v = a1.w;
                    tr.cur_inst = (int32_t)fbits(a2.x); tr.cur = fbits(a2.y); tr.la = fbits(a2.z); tr.lb = fbits(a2.w);
                    const int sp = (int)fbits(a3.x);
                    tr.o = tr.cur_inst >= 0 ? object_ray(S.inst_trav[tr.cur_inst], tr.w) : tr.w;
                    tr.st.reset(lds, spill);
                    for (int i = 0; i < sp; i += 4) {
                        const float4 e = sl[TAIL_SLOT_HEAD + i / 4];
                        tr.st.set(i, fbits(e.x));
                        if (i + 1 < sp) tr.st.set(i + 1, fbits(e.y));
                        if (i + 2 < sp) tr.st.set(i + 2, fbits(e.z));
                        if (i + 3 < sp) tr.st.set(i + 3, fbits(e.w));
                    }
                    tr.st.sp = sp;
                } else {
                    if (COUNT) ++n_vertices;
                    seed = seed + (uint32_t)ps.depth;
                    busy = tr.init(S, ps.ray, 0.0f, __builtin_inff(), lds, spill);
                }
                const uint32_t b0 = COUNT ? c.interior + c.leaf : 0u, t0 = COUNT ? c.tlas : 0u;
                // TAIL_CHECK steps per lane between looks at how many lanes still trace
                for (;;) {
                    for (int j = 0; j < TAIL_CHECK && busy; ++j) busy = tr.template step<COUNT>(S, &c);
                    const uint64_t m = __ballot(busy);   // the same for every lane in here
                    if (m == 0ull) break;
                    if (pool && !drained && __popcll(m) < suspend && __popcll(m) <= TAIL_POOL - npool &&
                        __ballot(busy && tr.st.sp > TAIL_SAVE_STACK) == 0ull)
                        break;
                }
                if (COUNT && heat_px) {
                    atomicAdd(&heat_px[0], (int32_t)(c.interior + c.leaf - b0));
                    atomicAdd(&heat_px[1], (int32_t)(c.tlas - t0));
                }
                // the traces still running are parked in the wave's slots from npool on
                const uint64_t mp = __ballot(busy);
                if (busy) {
                    // the path as a queued one (seed already advanced: its trace has begun); the
                    // pixel in the origin's w stays
                    *reinterpret_cast<float3*>(&W.ray_o[p]) = make_float3(ps.ray.o.x, ps.ray.o.y, ps.ray.o.z);
                    W.ray_d[p] = make_float4(ps.ray.d.x, ps.ray.d.y, ps.ray.d.z, bitsf(seed));
                    W.beta[p] = make_float4(ps.beta.x, ps.beta.y, ps.beta.z, ps.p_b);
                    W.L[p] = make_float4(ps.L.x, ps.L.y, ps.L.z,
                                         bitsf((uint32_t)ps.depth | (ps.any_nonspec ? 0x10000u : 0u) |
                                               (ps.wide_in ? 0x20000u : 0u) | (ps.cached ? 0x40000u : 0u)));
                    W.prev_pos[p] = make_float4(ps.prev_pos.x, ps.prev_pos.y, ps.prev_pos.z, 0.0f);
                    float4* sl = pool + (npool + __popcll(mp & lanes_below())) * TAIL_SLOT_F4;
                    const int sp = tr.st.sp;
                    sl[0] = make_float4(bitsf((uint32_t)p), tr.tmin, tr.tmax, tr.h.t);
                    sl[1] = make_float4(bitsf((uint32_t)tr.h.slot), bitsf((uint32_t)tr.h.inst), tr.h.u, tr.h.v);
                    sl[2] = make_float4(bitsf((uint32_t)tr.cur_inst), bitsf(tr.cur), bitsf(tr.la), bitsf(tr.lb));
                    sl[3] = make_float4(bitsf((uint32_t)sp), 0.0f, 0.0f, 0.0f);
                    for (int i = 0; i < sp; i += 4) {
                        const uint32_t e0 = tr.st.at(i);
                        const uint32_t e1 = i + 1 < sp ? tr.st.at(i + 1) : 0u;
                        const uint32_t e2 = i + 2 < sp ? tr.st.at(i + 2) : 0u;
                        const uint32_t e3 = i + 3 < sp ? tr.st.at(i + 3) : 0u;
                        sl[TAIL_SLOT_HEAD + i / 4] = make_float4(bitsf(e0), bitsf(e1), bitsf(e2), bitsf(e3));
                    }
                    parked = true;
                } else {
                    o = tr.st.overflow;
                    more = li_shade<COUNT, true, CACHE>(S, fc, ps, seed, tr.h, lds, &o, &vs, &c, &missed, &cb, &cf,
                                                        p, spill);
                }
            }
            ovf = ovf || o;
            if (parked) {
                active = false;
            } else if (!more || is_zero(ps.beta)) {
                W.L[p] = make_float4(ps.L.x, ps.L.y, ps.L.z, bitsf((uint32_t)ps.depth | (st & 0x10000u)));
                active = false;
            }
        }
        npool += __popcll(__ballot(parked));   // wave-uniform
        // wave-uniform: one atomic per wave for the lanes whose ray escaped
        const uint64_t mm = __ballot(missed);
        if (mm != 0ull) {
            const int base = wave_claim(&W.counters[cs(C_SHADOW)], __popcll(mm));
            if (missed) {
                const int k = base + __popcll(mm & lanes_below());
                W.shadow_queue[k] = p;
                W.shadow_d[k] = make_float4(ps.ray.d.x, ps.ray.d.y, ps.ray.d.z, 0.0f);
                W.shadow_c[k] = make_float4(ps.beta.x, ps.beta.y, ps.beta.z, 0.0f);
            }
        }
    }
    if (ovf) atomicOr(overflow, 1);
    if (COUNT) {
        uint32_t v[7] = {c.interior, c.leaf, c.tri, c.entries, c.tlas, n_paths, n_vertices};
        for (int k = 0; k < 7; ++k) {
            unsigned long long x = v[k];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
            if ((threadIdx.x & 63) == 0) atomicAdd(&counters[8 + k], x);
        }
    }
}

// Sky term of every path the tail ended by a miss: L + sky * beta, the same
// operations li_bounce makes at the miss, with all lanes of a wave busy.
__global__ __launch_bounds__(WF_BLOCK) void k_sky(FrameConst fc, WavefrontState W) {
    const int i = (int)(blockIdx.x * WF_BLOCK + threadIdx.x);
    if (i >= W.counters[cs(C_SHADOW)]) return;
    const int p = W.shadow_queue[i];
    const float4 d = W.shadow_d[i], b = W.shadow_c[i], l = W.L[p];
    const v3 sky = miss_radiance(fc, mk3(d.x, d.y, d.z), false);
    const v3 L = mk3(l.x, l.y, l.z) + sky * mk3(b.x, b.y, b.z);
    W.L[p] = make_float4(L.x, L.y, L.z, l.w);
}

// Visibility of the deferred next-event estimates; adds their term to L.  An
// estimate made at a cache-training vertex (record index in shadow_d.w, -1
// for none) also moves that record's radiance-so-far past itself.
__device__ __forceinline__ void shadow_resolve(WavefrontState& W, float4* t_lbefore, int k, bool occluded) {
    const int p = W.shadow_queue[k];
    const float4 c = occluded ? W.shadow_z[k] : W.shadow_c[k];
    const float4 l = W.L[p];
    const float4 nl = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w);
    W.L[p] = nl;
    const int32_t rec = (int32_t)fbits(W.shadow_d[k].w);
    if (rec >= 0) t_lbefore[rec] = make_float4(nl.x, nl.y, nl.z, 0.0f);
}

__global__ __launch_bounds__(WF_BLOCK) KPT_WPE_TRACE void k_shadow(SceneView S, WavefrontState W, int pull_slot, int32_t* overflow,
                                                     TraceTuning tune, const int32_t* order, float4* t_lbefore) {
    extern __shared__ uint32_t s_dyn[];   // tune.stack x blockDim.x words
    uint32_t spill[SPILL_STACK];
    const int count = W.counters[cs(C_SHADOW)];
    int32_t* ctr = &W.counters[cs(pull_slot)];
    const StackShape lds{s_dyn + threadIdx.x, (int)blockDim.x, tune.stack};
    bool ovf = false;
    if (!tune.refill) {
        const int chunk = (tune.chunk / WAVE) * WAVE;
        for (;;) {
            const int base = wave_claim(ctr, chunk);
            if (base >= count) break;
            for (int off = 0; off < chunk && base + off < count; off += WAVE) {
                const int kq = base + off + lane_id();
                if (kq >= count) continue;
                const int k = order ? order[kq] : kq;
                const float4 o = W.shadow_o[k], d = W.shadow_d[k];
                AnyTrav tr;
                if (tr.init(S, make_ray(mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z)), 0.0f, o.w, lds, spill)) {
                    while (tr.template step<false>(S, nullptr)) {}
                }
                ovf = ovf || tr.st.overflow;
                shadow_resolve(W, t_lbefore, k, tr.hit);
            }
        }
    } else {
        WaveFeeder feeder;
        feeder.init(tune);
        AnyTrav tr;
        int k = -1;
        bool busy = false;
        for (;;) {
            const int kk = feeder.feed(ctr, count, !busy);
            bool done_now = false;
            if (kk >= 0) {
                k = order ? order[kk] : kk;
                const float4 o = W.shadow_o[k], d = W.shadow_d[k];
                busy = tr.init(S, make_ray(mk3(o.x, o.y, o.z), mk3(d.x, d.y, d.z)), 0.0f, o.w, lds, spill);
                done_now = !busy;
            }
            if (!__any(busy || done_now)) {   // wave-uniform
                if (feeder.exhausted()) break;
                continue;
            }
            if (busy) {
                busy = tr.template step<false>(S, nullptr);
                done_now = !busy;
            }
            if (done_now) {
                ovf = ovf || tr.st.overflow;
                shadow_resolve(W, t_lbefore, k, tr.hit);
            }
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
    write_pixel(S, fc, fb, frame, i, (int)(i % fc.width), (int)(i / fc.width), L, mk3(va.x, va.y, va.z),
                mk3(vn.x, vn.y, vn.z), va.w, inst, prim, mk3(vp.x, vp.y, vp.z));
}

int g_num_cus = 0;

}  // namespace

hipError_t wavefront_frame(hipStream_t s, const SceneView& S, const FrameConst& fc, int32_t frame,
                           const FrameBuffers& fb, WavefrontState& W, int32_t* overflow, int32_t num_materials,
                           int32_t sort_by_material, const TraceTuning& tune, WfInstrument* ins,
                           const CacheHooks* cache, const ViewBatch* views, hipEvent_t finish_after) {
    if (ins) {
        ins->launches = 0;
        ins->tail_launches = 0;
        for (int x = 0; x < WF_XSTAGES; ++x) ins->xlaunches[x] = 0;
    }
    CacheBuffers cb{};
    CacheFrame cf{};
    // the cache frame starts here, or (finish_after given) once the previous
    // frame is done, before this frame's first launch that touches the cache
    bool cache_started = true;
    auto start_cache = [&]() -> hipError_t {
        if (cache_started) return hipSuccess;
        cache_started = true;
        if (finish_after) {
            const hipError_t e = hipStreamWaitEvent(s, finish_after, 0);
            if (e != hipSuccess) return e;
        }
        return launch_cache_frame_start(s, cb, cf);
    };
    if (cache && cache->cb && (cache->cf.enabled || cache->cf.train)) {
        cb = *cache->cb;
        cf = cache->cf;
        cache_started = false;
        if (!finish_after) {
            hipError_t e = start_cache();
            if (e != hipSuccess) return e;
        }
    }
    if (g_num_cus == 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) g_num_cus = prop.multiProcessorCount;
        if (g_num_cus <= 0) g_num_cus = 256;
    }
    const int64_t npx = views ? views->total : (int64_t)fc.width * fc.height;
    const unsigned gblocks = (unsigned)((npx + WF_BLOCK - 1) / WF_BLOCK);
    // the per-pixel DebugData: counted single-view frames only, cleared per frame
    if (views || !(ins && ins->counters)) W.heat = nullptr;
    if (W.heat) {
        const int64_t nh = 2 * npx;
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, W.heat, 0, (int)nh);
    }
    // persistent grids: blocks_per_cu workgroups of 256 lanes per CU; the
    // trace-bearing kernels run the same number of lanes in workgroups of
    // tune.block lanes, each with tune.stack LDS stack entries per lane
    const unsigned pgrid = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)g_num_cus * std::max(1, tune.blocks_per_cu), gblocks));
    const int tblock = tune.block;
    const unsigned tgrid = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)g_num_cus * std::max(1, tune.blocks_per_cu) * (WF_BLOCK / tblock),
                             (npx + tblock - 1) / tblock));
    const size_t tlds = (size_t)tune.stack * (size_t)tblock * sizeof(uint32_t);
    const int nkeys = std::min(MAX_KEYS, num_materials + 1);
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((WF_COUNTER_WORDS + 255) / 256)), dim3(256), 0, s, W.counters, 0,
                       WF_COUNTER_WORDS);
    if (views)
        hipLaunchKernelGGL(k_generate_views, dim3(gblocks), dim3(WF_BLOCK), 0, s, fc, frame, W, *views);
    else
        hipLaunchKernelGGL(k_generate, dim3(gblocks), dim3(WF_BLOCK), 0, s, fc, frame, W);
    int max_bounces = std::min(fc.max_depth + 1, MAX_BOUNCE_SLOTS);
    // When the cache renders and every lobe leads to a cache vertex, no path
    // passes vertex 2 (others end at vertex 1, training paths at their second
    // cache vertex): the launches for deeper bounces would find empty queues.
    if (cf.enabled && !(cf.wide_roughness > 0.0f)) max_bounces = std::min(max_bounces, 3);
    for (int bounce = 0; bounce < max_bounces; ++bounce) {
        const int parity = bounce & 1;
        if (bounce == 1) {
            const hipError_t e = start_cache();
            if (e != hipSuccess) return e;
        }
        if (tune.tail > 0 && bounce >= tune.tail) {
            // the tail's paths in spatial order of their rays (ray_sort knob)
            const int32_t* tail_q = nullptr;
            if (tune.ray_sort && W.rtemp) {
                hipLaunchKernelGGL(k_ray_keys, dim3(gblocks), dim3(WF_BLOCK), 0, s, fc, W, parity, npx);
                size_t bytes = W.rtemp_bytes;
                hipError_t e = rocprim::radix_sort_pairs(W.rtemp, bytes, W.rkeys[0], W.rkeys[1], W.rvals[0],
                                                         W.rvals[1], (size_t)npx, 0, RKEY_BITS, s);
                if (e != hipSuccess) return e;
                tail_q = W.rvals[1];
            }
            hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, W.counters + cs(C_SHADOW), 0, 1);
            if (ins) (void)hipEventRecord(ins->tail_begin, s);
            const bool with_cache = cf.enabled || cf.train;
            unsigned long long* tc = (ins && ins->counters) ? ins->counters : nullptr;
            if (tc && with_cache)
                hipLaunchKernelGGL((k_tail<true, true>), dim3(tgrid), dim3(tblock), tlds, s, S, fc, W, parity,
                                   C_PULL + 3 * bounce, overflow, tc, tail_q, cb, cf, tune.stack, tune.suspend);
            else if (tc)
                hipLaunchKernelGGL((k_tail<true, false>), dim3(tgrid), dim3(tblock), tlds, s, S, fc, W, parity,
                                   C_PULL + 3 * bounce, overflow, tc, tail_q, cb, cf, tune.stack, tune.suspend);
            else if (with_cache)
                hipLaunchKernelGGL((k_tail<false, true>), dim3(tgrid), dim3(tblock), tlds, s, S, fc, W, parity,
                                   C_PULL + 3 * bounce, overflow, tc, tail_q, cb, cf, tune.stack, tune.suspend);
            else
                hipLaunchKernelGGL((k_tail<false, false>), dim3(tgrid), dim3(tblock), tlds, s, S, fc, W, parity,
                                   C_PULL + 3 * bounce, overflow, tc, tail_q, cb, cf, tune.stack, tune.suspend);
            if (ins) {
                (void)hipEventRecord(ins->tail_end, s);
                ins->tail_launches = 1;
            }
            if (ins) (void)hipEventRecord(ins->xbegin[WF_X_SKY], s);
            hipLaunchKernelGGL(k_sky, dim3(gblocks), dim3(WF_BLOCK), 0, s, fc, W);
            if (ins) {
                (void)hipEventRecord(ins->xend[WF_X_SKY], s);
                ins->xlaunches[WF_X_SKY] = 1;
            }
            break;
        }
        // next queue and shadow count start at zero
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, W.counters + cs(C_QCOUNT + (parity ^ 1)), 0, 1);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, W.counters + cs(C_SHADOW), 0, 1);
        // continuation rays (not the pixel-ordered camera rays) in spatial order
        const int32_t* trace_q = nullptr;
        if (tune.ray_sort && bounce > 0 && W.rtemp) {
            hipLaunchKernelGGL(k_ray_keys, dim3(gblocks), dim3(WF_BLOCK), 0, s, fc, W, parity, npx);
            size_t bytes = W.rtemp_bytes;
            hipError_t e = rocprim::radix_sort_pairs(W.rtemp, bytes, W.rkeys[0], W.rkeys[1], W.rvals[0], W.rvals[1],
                                                     (size_t)npx, 0, RKEY_BITS, s);
            if (e != hipSuccess) return e;
            trace_q = W.rvals[1];
        }
        const bool timed = ins && ins->launches < WF_MAX_TRACE_LAUNCHES;
        if (timed) (void)hipEventRecord(ins->begin[ins->launches], s);
        if (ins && ins->counters) {
            hipLaunchKernelGGL(k_trace<true>, dim3(tgrid), dim3(tblock), tlds, s, S, W, parity, C_PULL + 3 * bounce,
                               overflow, ins->counters, tune, trace_q);
        } else {
            hipLaunchKernelGGL(k_trace<false>, dim3(tgrid), dim3(tblock), tlds, s, S, W, parity, C_PULL + 3 * bounce,
                               overflow, (unsigned long long*)nullptr, tune, trace_q);
        }
        if (timed) (void)hipEventRecord(ins->end[ins->launches++], s);
        const int32_t* shade_q = W.queue[parity];
        // sort 2: the primary hits keep their tile order (their shadow rays and
        // the bounce-1 rays then leave from neighbouring points)
        if (sort_by_material && nkeys > 1 && !(sort_by_material == 2 && bounce == 0)) {
            hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(2 * MAX_KEYS), 0, s, W.hist, 0, 2 * MAX_KEYS);
            hipLaunchKernelGGL(k_sort_hist, dim3(pgrid), dim3(WF_BLOCK), 0, s, S, W, parity, nkeys);
            hipLaunchKernelGGL(k_sort_offsets, dim3(1), dim3(64), 0, s, W, nkeys);
            hipLaunchKernelGGL(k_sort_scatter, dim3(pgrid), dim3(WF_BLOCK), 0, s, S, W, parity, nkeys);
            shade_q = W.sort_tmp;
        }
        const bool time_shade = ins && bounce == 0;
        if (time_shade) (void)hipEventRecord(ins->xbegin[WF_X_SHADE0], s);
        hipLaunchKernelGGL(k_shade, dim3(gblocks), dim3(WF_BLOCK), 0, s, S, fc, W, parity, shade_q, cb, cf);
        if (time_shade) {
            (void)hipEventRecord(ins->xend[WF_X_SHADE0], s);
            ins->xlaunches[WF_X_SHADE0] = 1;
        }
        const int32_t* shadow_order = nullptr;
        if (tune.ray_sort && W.rtemp) {
            hipLaunchKernelGGL(k_shadow_keys, dim3(gblocks), dim3(WF_BLOCK), 0, s, fc, W, npx);
            size_t bytes = W.rtemp_bytes;
            hipError_t e = rocprim::radix_sort_pairs(W.rtemp, bytes, W.rkeys[0], W.rkeys[1], W.rvals[0], W.rvals[1],
                                                     (size_t)npx, 0, RKEY_BITS, s);
            if (e != hipSuccess) return e;
            shadow_order = W.rvals[1];
        }
        hipLaunchKernelGGL(k_shadow, dim3(tgrid), dim3(tblock), tlds, s, S, W, C_PULL + 3 * bounce + 2, overflow,
                           tune, shadow_order, cb.t_lbefore);
    }
    {
        const hipError_t e = start_cache();   // frames that end at bounce 0
        if (e != hipSuccess) return e;
    }
    if (cf.enabled) {
        if (ins) (void)hipEventRecord(ins->xbegin[WF_X_INFER], s);
        hipError_t e = launch_cache_infer_paths(s, cb, cache->enc, W.L, npx, cache->num_cus);
        if (e != hipSuccess) return e;
        if (ins) {
            (void)hipEventRecord(ins->xend[WF_X_INFER], s);
            ins->xlaunches[WF_X_INFER] = 1;
        }
    }
    if (cf.train) {
        hipError_t e = launch_cache_targets(s, cb, W.L, cache->enc.scale);
        if (e != hipSuccess) return e;
    }
    if (finish_after) {
        const hipError_t e = hipStreamWaitEvent(s, finish_after, 0);
        if (e != hipSuccess) return e;
    }
    if (views)
        hipLaunchKernelGGL(k_finish_views, dim3(gblocks), dim3(WF_BLOCK), 0, s, W, *views);
    else
        hipLaunchKernelGGL(k_finish, dim3(gblocks), dim3(WF_BLOCK), 0, s, S, fc, frame, W, fb);
    return hipGetLastError();
}

size_t ray_sort_temp_bytes(int64_t n) {
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (int32_t*)nullptr,
                                  (int32_t*)nullptr, (size_t)n, 0, RKEY_BITS) != hipSuccess)
        return 0;
    return bytes;
}

hipError_t launch_views_mean(hipStream_t s, float4* out, int64_t n, int32_t spp) {
    if (n <= 0 || spp <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_views_mean, dim3((unsigned)((n + WF_BLOCK - 1) / WF_BLOCK)), dim3(WF_BLOCK), 0, s, out, n,
                       spp);
    return hipGetLastError();
}

hipError_t launch_heat_colour(hipStream_t s, const int32_t* heat, int64_t n, float4* rgba) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_heat_colour, dim3((unsigned)((n + WF_BLOCK - 1) / WF_BLOCK)), dim3(WF_BLOCK), 0, s, heat, n,
                       rgba);
    return hipGetLastError();
}

}  // namespace kpt
