// kpt_cache.hip — the radiance-cache MLP on gfx950 matrix cores.
//
// The reference renderer has no radiance cache; this is the learned part of
// the path tracer: a 6-layer, 64-wide ReLU MLP over a 32-feature frequency
// encoding of (position, direction, roughness), BF16 weights and activations
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation, trained online with
// Adam on the tracer's own path suffixes.
//
// MFMA operand layout (16x16x32, wave of 64 lanes, r = lane % 16,
// g = lane / 16): lane holds A[r][8g .. 8g+7], B[8g .. 8g+7][r] and
// D[4g .. 4g+3][r].  Every GEMM is arranged so the lane's 8 K-values are
// contiguous in memory (one 16-byte load):
//   forward        Y^T[out][q]  = W[out][in]      * X^T[in][q]   (X row-major [q][in])
//   input grad     dX^T[in][q]  = W^T[in][out]    * dY^T[out][q] (W^T kept as its own copy)
//   weight grad    dW[out][in]  = dY^T[out][q]    * X[q][in]     (X^T and dY^T kept in LDS)
// so D hands lane (r, g) the outputs 4g..4g+3 of query r of each 16-wide tile.
//
// Kernels
//   k_rc_infer<MODE>  inference over queries: MODE 0 adds throughput * radiance
//                     to the path's radiance, MODE 1 writes the raw outputs.
//                     Weights (40 KB BF16) staged once per workgroup in LDS,
//                     one 16-query tile per wave, activations in a per-wave
//                     LDS row buffer.
//   k_rc_train        forward + relative-L2 loss + backward for 32 samples per
//                     workgroup (2 waves); writes its weight-gradient partial.
//   k_rc_reduce       sums consecutive chunks of the partials (workgroup order);
//   k_rc_adam         adds the chunk sums in chunk order (deterministic), Adam
//                     step on the FP32 master weights, BF16 republish.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../common/kpt_dmath.h"
#include "kpt_cache.h"
#include "kpt_launch.h"

namespace kpt {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float RC_PI = 3.14159265358979323846f;
constexpr float RC_HALF_PI = 1.57079632679489661923f;
constexpr int TRAIN_Q = 32;                 // samples per training workgroup
constexpr int TRAIN_BLOCK = 128;            // 2 waves of 16 samples

// float -> bfloat16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ bf16x8 pack8(const uint16_t* h) {
    uint4 u;
    u.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    u.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
    u.z = (uint32_t)h[4] | ((uint32_t)h[5] << 16);
    u.w = (uint32_t)h[6] | ((uint32_t)h[7] << 16);
    return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 zero4() { f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f}; return z; }
// lanes of one wave exchange data through LDS: order their accesses
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float comp(float4 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// Input feature idx of a query (position xyz + roughness in pos, outgoing
// direction in dir): sin/cos of the normalised position at 1, 2, 4, 8 cycles
// per box (24), sin/cos of (pi/2) * direction (6), roughness, constant 1.
__device__ __forceinline__ float rc_feature(int idx, float4 pos, float4 dir, const CacheEncode& e) {
    if (idx < 24) {
        const int k = idx / 6, a = (idx % 6) >> 1;
        const float ph = (comp(pos, a) - e.bmin[a]) * e.binv[a];
        const float arg = (RC_PI * (float)(1 << k)) * ph;
        return (idx & 1) ? dm::cosf(arg) : dm::sinf(arg);
    }
    if (idx < 30) {
        const int a = (idx - 24) >> 1;
        const float arg = RC_HALF_PI * comp(dir, a);
        return (idx & 1) ? dm::cosf(arg) : dm::sinf(arg);
    }
    return idx == 30 ? pos.w : 1.0f;
}

// the lane's 8 features 8g..8g+7 as one B operand
__device__ __forceinline__ bf16x8 rc_features8(int g, float4 pos, float4 dir, const CacheEncode& e, uint16_t* h) {
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = f2bf(rc_feature(8 * g + j, pos, dir, e));
    return pack8(h);
}

// ReLU + BF16 of one 16-wide output tile into a row-major [q][64] buffer
__device__ __forceinline__ void store_relu_row(uint16_t* row, int t, int g, f32x4 y) {
    uint2 v;
    v.x = (uint32_t)f2bf(fmaxf(y[0], 0.0f)) | ((uint32_t)f2bf(fmaxf(y[1], 0.0f)) << 16);
    v.y = (uint32_t)f2bf(fmaxf(y[2], 0.0f)) | ((uint32_t)f2bf(fmaxf(y[3], 0.0f)) << 16);
    *reinterpret_cast<uint2*>(row + 16 * t + 4 * g) = v;
}

// ------------------------------------------------------------------ inference
template <int MODE>
__global__ __launch_bounds__(256) void k_rc_infer(CacheBuffers cb, CacheEncode enc, const float4* qpos,
                                                  const float4* qdir, const float4* qbeta, const int32_t* count_dev,
                                                  int64_t count_host, int64_t cap, float4* dst) {
    __shared__ __attribute__((aligned(16))) uint16_t s_w[RC_PADDED];
    __shared__ __attribute__((aligned(16))) uint16_t s_act[4][16 * RC_HID];
    for (int i = threadIdx.x; i < RC_PADDED / 8; i += 256)
        reinterpret_cast<uint4*>(s_w)[i] = reinterpret_cast<const uint4*>(cb.wb)[i];
    __syncthreads();
    const int64_t n = count_dev ? std::min<int64_t>((int64_t)*count_dev, cap) : count_host;
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63), r = lane & 15, g = lane >> 4;
    uint16_t* act = s_act[wave];
    uint16_t* row = act + r * RC_HID;
    const int64_t tiles = (n + 15) / 16;
    const int64_t stride = (int64_t)gridDim.x * 4;
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // the next tile's query (and, adding into paths, its throughput and the
    // path's radiance) is loaded while the current tile runs its layers
    int64_t tile = (int64_t)blockIdx.x * 4 + wave;
    float4 p4n = z4, d4n = z4, ben = z4, lpn = z4;
    if (tile < tiles && tile * 16 + r < n) {
        p4n = qpos[tile * 16 + r];
        d4n = qdir[tile * 16 + r];
        if (MODE == 0) {
            ben = qbeta[tile * 16 + r];
            lpn = dst[(int)__builtin_bit_cast(uint32_t, d4n.w)];
        }
    }
    for (; tile < tiles; tile += stride) {
        const int64_t qi = tile * 16 + r;
        const bool valid = qi < n;
        const float4 p4 = p4n, d4 = d4n, be = ben, lp = lpn;
        const int64_t ni = (tile + stride) * 16 + r;
        if (tile + stride < tiles && ni < n) {
            p4n = qpos[ni];
            d4n = qdir[ni];
            if (MODE == 0) {
                ben = qbeta[ni];
                lpn = dst[(int)__builtin_bit_cast(uint32_t, d4n.w)];
            }
        }
        uint16_t h[8];
        const bf16x8 x = rc_features8(g, p4, d4, enc, h);
        // layer 0 (K = 32)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 y = mfma(ld8(s_w + rc_off(0) + (16 * t + r) * RC_IN + 8 * g), x, zero4());
            store_relu_row(row, t, g, y);
        }
        wave_lds_sync();
        // layers 1..4 (K = 64): the whole input row is loaded before it is overwritten
#pragma unroll
        for (int L = 1; L <= 4; ++L) {
            const bf16x8 b0 = ld8(row + 8 * g), b1 = ld8(row + 32 + 8 * g);
            wave_lds_sync();
            const uint16_t* wl = s_w + rc_off(L);
            f32x4 y[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                y[t] = mfma(ld8(wl + (16 * t + r) * RC_HID + 8 * g), b0, zero4());
                y[t] = mfma(ld8(wl + (16 * t + r) * RC_HID + 32 + 8 * g), b1, y[t]);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) store_relu_row(row, t, g, y[t]);
            wave_lds_sync();
        }
        // output layer: rows 0..15 of the 32-row padded block
        const bf16x8 b0 = ld8(row + 8 * g), b1 = ld8(row + 32 + 8 * g);
        const uint16_t* w5 = s_w + rc_off(5);
        f32x4 o = mfma(ld8(w5 + r * RC_HID + 8 * g), b0, zero4());
        o = mfma(ld8(w5 + r * RC_HID + 32 + 8 * g), b1, o);
        wave_lds_sync();
        if (valid && g == 0) {
            if (MODE == 0) {
                const int p = (int)__builtin_bit_cast(uint32_t, d4.w);
                dst[p] = make_float4(lp.x + be.x * (enc.scale * fmaxf(o[0], 0.0f)),
                                     lp.y + be.y * (enc.scale * fmaxf(o[1], 0.0f)),
                                     lp.z + be.z * (enc.scale * fmaxf(o[2], 0.0f)), lp.w);
            } else {
                dst[qi] = make_float4(o[0], o[1], o[2], 0.0f);
            }
        }
    }
}

// ------------------------------------------------------------------ training
__host__ __device__ constexpr int xo(int l) { return l == 0 ? 0 : TRAIN_Q * RC_IN + (l - 1) * TRAIN_Q * RC_HID; }
constexpr int XS = TRAIN_Q * RC_IN + 5 * TRAIN_Q * RC_HID;   // 11264 BF16 per copy

__global__ __launch_bounds__(TRAIN_BLOCK) void k_rc_train(CacheBuffers cb, CacheEncode enc, const float4* qpos,
                                                          const float4* qdir, const float4* target,
                                                          const int32_t* count_dev, int64_t count_host, int64_t cap,
                                                          int64_t first, int32_t batch) {
    __shared__ __attribute__((aligned(16))) uint16_t sX[XS];            // layer inputs [q][in]
    __shared__ __attribute__((aligned(16))) uint16_t sXT[XS];           // [in][q]
    __shared__ __attribute__((aligned(16))) uint16_t sdY[TRAIN_Q * RC_HID];   // [q][out]
    __shared__ __attribute__((aligned(16))) uint16_t sdYT[RC_HID * TRAIN_Q];  // [out][q]
    __shared__ float s_loss[2];
    const int64_t n = count_dev ? std::min<int64_t>((int64_t)*count_dev, cap) : count_host;
    const int64_t last = std::min<int64_t>(first + batch, n);
    const int64_t b_eff = last - first;
    if (b_eff <= 0) return;
    const float inv_norm = 1.0f / (3.0f * (float)b_eff);
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63), r = lane & 15, g = lane >> 4;
    const int q = wave * 16 + r;
    const int64_t qi = first + (int64_t)blockIdx.x * TRAIN_Q + q;
    const bool valid = qi < last;
    float* part = cb.partial + (size_t)blockIdx.x * RC_PADDED;

    // ---- forward, keeping every layer input (and its transpose) in LDS
    const float4 p4 = valid ? qpos[qi] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 d4 = valid ? qdir[qi] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint16_t h[8];
    const bf16x8 x = rc_features8(g, p4, d4, enc, h);
    *reinterpret_cast<bf16x8*>(sX + xo(0) + q * RC_IN + 8 * g) = x;
#pragma unroll
    for (int j = 0; j < 8; ++j) sXT[xo(0) + (8 * g + j) * TRAIN_Q + q] = h[j];
    f32x4 y[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) y[t] = mfma(ld8(cb.wb + rc_off(0) + (16 * t + r) * RC_IN + 8 * g), x, zero4());
    for (int L = 1; L <= 5; ++L) {
        // y holds layer L-1's pre-activations: ReLU -> layer L input
        uint16_t* xr = sX + xo(L) + q * RC_HID;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint16_t v = f2bf(fmaxf(y[t][j], 0.0f));
                xr[16 * t + 4 * g + j] = v;
                sXT[xo(L) + (16 * t + 4 * g + j) * TRAIN_Q + q] = v;
            }
        }
        wave_lds_sync();
        const bf16x8 b0 = ld8(xr + 8 * g), b1 = ld8(xr + 32 + 8 * g);
        const uint16_t* wl = cb.wb + rc_off(L);
        const int tiles = (L == 5) ? 1 : 4;
        for (int t = 0; t < tiles; ++t) {
            y[t] = mfma(ld8(wl + (16 * t + r) * RC_HID + 8 * g), b0, zero4());
            y[t] = mfma(ld8(wl + (16 * t + r) * RC_HID + 32 + 8 * g), b1, y[t]);
        }
    }
    // ---- relative L2 loss (o - t)^2 / (o^2 + 0.01), denominator held constant
    float loss = 0.0f;
    float dy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (g == 0 && valid) {
        const float4 t4 = target[qi];
        const float tt[3] = {t4.x, t4.y, t4.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float o = y[0][c];
            const float d = o - tt[c];
            const float den = o * o + 0.01f;
            loss += d * d / den;
            dy[c] = (2.0f * d / den) * inv_norm;
        }
    }
    // output-layer gradient, padded to 32 columns
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint16_t v = f2bf(g == 0 ? dy[j] : 0.0f);
        sdY[q * RC_HID + 4 * g + j] = v;
        sdY[q * RC_HID + 16 + 4 * g + j] = 0;
        sdYT[(4 * g + j) * TRAIN_Q + q] = v;
        sdYT[(16 + 4 * g + j) * TRAIN_Q + q] = 0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (lane == 0) s_loss[wave] = loss;
    __syncthreads();

    // ---- backward
    for (int L = 5; L >= 0; --L) {
        const int in = rc_in(L), opad = rc_out_pad(L);
        const int ntn = in / 16, ntiles = (opad / 16) * ntn;
        // weight gradient tiles over the workgroup's 32 samples (K = 32)
        for (int ti = wave; ti < ntiles; ti += 2) {
            const int tm = ti / ntn, tn = ti % ntn;
            const f32x4 acc = mfma(ld8(sdYT + (16 * tm + r) * TRAIN_Q + 8 * g),
                                   ld8(sXT + xo(L) + (16 * tn + r) * TRAIN_Q + 8 * g), zero4());
#pragma unroll
            for (int j = 0; j < 4; ++j) part[rc_off(L) + (16 * tm + 4 * g + j) * in + 16 * tn + r] = acc[j];
        }
        if (L == 0) break;
        // input gradient of the wave's own 16 samples, masked by the ReLU of layer L-1
        f32x4 dx[4];
        const uint16_t* wt = cb.wbt + rc_off(L);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = mfma(ld8(wt + (16 * t + r) * opad + 8 * g), ld8(sdY + q * RC_HID + 8 * g), zero4());
            if (opad == 64) acc = mfma(ld8(wt + (16 * t + r) * opad + 32 + 8 * g), ld8(sdY + q * RC_HID + 32 + 8 * g), acc);
            dx[t] = acc;
        }
        __syncthreads();   // every wave is done reading layer L's dY
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int col = 16 * t + 4 * g + j;
                const float xv = bf2f(sX[xo(L) + q * RC_HID + col]);
                const uint16_t v = f2bf(xv > 0.0f ? dx[t][j] : 0.0f);
                sdY[q * RC_HID + col] = v;
                sdYT[col * TRAIN_Q + q] = v;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) cb.loss_partial[blockIdx.x] = s_loss[0] + s_loss[1];
}

// position of padded parameter i in the transposed [in][out_pad] copy
__device__ __forceinline__ int rc_transposed(int i) {
    int L = 5;
    while (L > 0 && i < rc_off(L)) --L;
    const int local = i - rc_off(L), in = rc_in(L);
    return rc_off(L) + (local % in) * rc_out_pad(L) + local / in;
}

struct AdamParams {
    float lr, beta1, beta2, eps;
    float bc1, bc2;       // 1 - beta^t
};

// First level of the gradient sum: chunk c (blockIdx.y) of `per` consecutive
// workgroup partials, summed in workgroup order, for 256 parameters per
// workgroup.  Many chunks run at once; k_rc_adam then adds the chunk sums in
// chunk order, so the total is the same for every launch.
__global__ __launch_bounds__(256) void k_rc_reduce(CacheBuffers cb, int nwg, int per, const int32_t* count_dev,
                                                   int64_t count_host, int64_t cap, int64_t first) {
    const int64_t n = count_dev ? std::min<int64_t>((int64_t)*count_dev, cap) : count_host;
    if (n - first <= 0) return;
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= RC_PARAMS) return;
    const int k0 = (int)blockIdx.y * per, k1 = std::min(nwg, k0 + per);
    float g = 0.0f;
    for (int k = k0; k < k1; ++k) g += cb.partial[(size_t)k * RC_PADDED + i];
    cb.partial_sum[(size_t)blockIdx.y * RC_PADDED + i] = g;
}

__global__ __launch_bounds__(256) void k_rc_adam(CacheBuffers cb, int nwg, int nchunks, const int32_t* count_dev,
                                                 int64_t count_host, int64_t cap, int64_t first, int32_t batch,
                                                 AdamParams ap) {
    const int64_t n = count_dev ? std::min<int64_t>((int64_t)*count_dev, cap) : count_host;
    const int64_t b_eff = std::min<int64_t>(first + batch, n) - first;
    if (b_eff <= 0) return;
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i == 0) {
        float ls = 0.0f;
        for (int k = 0; k < nwg; ++k) ls += cb.loss_partial[k];
        cb.stats[0] = ls / (3.0f * (float)b_eff);
        cb.stats[1] = (float)b_eff;
    }
    if (i >= RC_PARAMS) return;
    float gsum = 0.0f;
    for (int c = 0; c < nchunks; ++c) gsum += cb.partial_sum[(size_t)c * RC_PADDED + i];
    cb.grad[i] = gsum;
    const float m = ap.beta1 * cb.adam_m[i] + (1.0f - ap.beta1) * gsum;
    const float v = ap.beta2 * cb.adam_v[i] + (1.0f - ap.beta2) * (gsum * gsum);
    cb.adam_m[i] = m;
    cb.adam_v[i] = v;
    const float mh = m / ap.bc1, vh = v / ap.bc2;
    const float w = cb.w[i] - ap.lr * mh / (__builtin_sqrtf(vh) + ap.eps);
    cb.w[i] = w;
    const uint16_t b = f2bf(w);
    cb.wb[i] = b;
    cb.wbt[rc_transposed(i)] = b;
}

// BF16 copies of the master weights (padding stays zero)
__global__ void k_rc_publish(CacheBuffers cb) {
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= RC_PADDED) return;
    const uint16_t b = (i < RC_PARAMS) ? f2bf(cb.w[i]) : (uint16_t)0;
    cb.wb[i] = b;
    cb.wbt[rc_transposed(i)] = b;
}

// suffix radiance of each training record in network units
__global__ __launch_bounds__(256) void k_rc_targets(CacheBuffers cb, const float4* L, float scale) {
    const int64_t n = std::min<int64_t>((int64_t)cb.counters[1], cb.train_cap);
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int p = (int)__builtin_bit_cast(uint32_t, cb.t_dir[i].w);
    const float4 l = L[p], b = cb.t_lbefore[i], be = cb.t_beta[i];
    auto suffix = [&](float lend, float lb, float bc) { return bc > 0.0f ? ((lend - lb) / bc) / scale : 0.0f; };
    cb.t_target[i] = make_float4(suffix(l.x, b.x, be.x), suffix(l.y, b.y, be.y), suffix(l.z, b.z, be.z), 0.0f);
}

// Start of a frame's cache records: this frame's training threshold into
// counters[2], then the query and record counts cleared.  With cf.adapt the
// previous frame's counts (same context and resolution) tell how many paths
// reached their cache vertex: when the cache rendered, each of them wrote a
// query or a record; when it only trained, the records were the threshold's
// share of them.  The threshold is then set so that about cf.train_samples
// paths train; otherwise cf.train_threshold (train_samples over the frame's
// paths) stands.
__global__ void k_rc_frame_start(CacheBuffers cb, CacheFrame cf) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t thr = cf.train_threshold;
    if (cf.adapt) {
        const double q = (double)cb.counters[0], r = (double)cb.counters[1];
        const double prev = (double)(uint32_t)cb.counters[2];
        const double reach = cf.enabled ? q + r : (prev > 0.0 ? r * 16777216.0 / prev : 0.0);
        if (reach >= 1.0) thr = (uint32_t)fmin(16777216.0, ceil(16777216.0 * (double)cf.train_samples / reach));
    }
    cb.counters[2] = (int32_t)thr;
    cb.counters[0] = 0;
    cb.counters[1] = 0;
}

}  // namespace

int cache_train_workgroups(int32_t batch) { return (batch + TRAIN_Q - 1) / TRAIN_Q; }

hipError_t launch_cache_frame_start(hipStream_t s, const CacheBuffers& cb, const CacheFrame& cf) {
    hipLaunchKernelGGL(k_rc_frame_start, dim3(1), dim3(64), 0, s, cb, cf);
    return hipGetLastError();
}

hipError_t launch_cache_publish(hipStream_t s, const CacheBuffers& cb) {
    hipLaunchKernelGGL(k_rc_publish, dim3((RC_PADDED + 255) / 256), dim3(256), 0, s, cb);
    return hipGetLastError();
}

// Workgroups of k_rc_infer resident per CU (its 48 KB of LDS allow 3): the
// persistent grid is sized to what runs at once, so no workgroup waits for
// another to finish before it starts its share of the tiles.
int infer_blocks_per_cu() {
    static int nb = 0;
    if (nb == 0) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, reinterpret_cast<const void*>(&k_rc_infer<0>), 256, 0) !=
                hipSuccess || v < 1)
            v = 1;
        nb = v;
    }
    return nb;
}

hipError_t launch_cache_infer_paths(hipStream_t s, const CacheBuffers& cb, const CacheEncode& enc, float4* L,
                                    int64_t max_queries, int num_cus) {
    const int64_t tiles = (max_queries + 15) / 16;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4,
                                                                           (int64_t)num_cus * infer_blocks_per_cu()));
    hipLaunchKernelGGL(k_rc_infer<0>, dim3(grid), dim3(256), 0, s, cb, enc, cb.q_pos, cb.q_dir, cb.q_beta,
                       cb.counters, (int64_t)0, cb.query_cap, L);
    return hipGetLastError();
}

hipError_t launch_cache_infer_raw(hipStream_t s, const CacheBuffers& cb, const CacheEncode& enc, const float4* pos,
                                  const float4* dir, int64_t n, float4* out, int num_cus) {
    if (n <= 0) return hipSuccess;
    const int64_t tiles = (n + 15) / 16;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4,
                                                                           (int64_t)num_cus * infer_blocks_per_cu()));
    hipLaunchKernelGGL(k_rc_infer<1>, dim3(grid), dim3(256), 0, s, cb, enc, pos, dir, (const float4*)nullptr,
                       (const int32_t*)nullptr, n, n, out);
    return hipGetLastError();
}

hipError_t launch_cache_targets(hipStream_t s, const CacheBuffers& cb, const float4* L, float scale) {
    const unsigned grid = (unsigned)((cb.train_cap + 255) / 256);
    hipLaunchKernelGGL(k_rc_targets, dim3(grid), dim3(256), 0, s, cb, L, scale);
    return hipGetLastError();
}

hipError_t launch_cache_train_step(hipStream_t s, const CacheBuffers& cb, const CacheEncode& enc, const float4* pos,
                                   const float4* dir, const float4* target, const int32_t* count_dev,
                                   int64_t count_host, int64_t cap, int64_t first, int32_t batch, float lr,
                                   int64_t step) {
    const int nwg = cache_train_workgroups(batch);
    if (nwg > cb.max_wg) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_rc_train, dim3(nwg), dim3(TRAIN_BLOCK), 0, s, cb, enc, pos, dir, target, count_dev,
                       count_host, cap, first, batch);
    AdamParams ap;
    ap.lr = lr;
    ap.beta1 = 0.9f;
    ap.beta2 = 0.99f;
    ap.eps = 1e-15f;
    ap.bc1 = (float)(1.0 - std::pow(0.9, (double)step));
    ap.bc2 = (float)(1.0 - std::pow(0.99, (double)step));
    const int nchunks = std::min(RC_CHUNKS, nwg);
    const int per = (nwg + nchunks - 1) / nchunks;
    hipLaunchKernelGGL(k_rc_reduce, dim3((RC_PARAMS + 255) / 256, (unsigned)((nwg + per - 1) / per)), dim3(256), 0, s,
                       cb, nwg, per, count_dev, count_host, cap, first);
    hipLaunchKernelGGL(k_rc_adam, dim3((RC_PARAMS + 255) / 256), dim3(256), 0, s, cb, nwg, (nwg + per - 1) / per,
                       count_dev, count_host, cap, first, batch, ap);
    return hipGetLastError();
}

}  // namespace kpt
