// kpt_cache.h — the radiance-cache network shared by the kernels and the
// host side of the C-ABI: sizes, parameter layouts and the buffers.
//
// Network: 32 input features -> 64 -> 64 -> 64 -> 64 -> 64 -> 3, ReLU after
// every hidden layer, no biases (18,624 weights).  Weights are stored per
// layer, row-major [out][in].  The BF16 copies used by the MFMA kernels pad
// the output layer to 32 rows (zeros); the transposed BF16 copy is [in][out].
// Real parameters are exactly the first RC_PARAMS entries of the padded
// layout, so one index addresses the FP32 master, the gradient and the BF16
// copy.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kpt {

constexpr int RC_IN = 32;
constexpr int RC_HID = 64;
constexpr int RC_OUT = 3;
constexpr int RC_LAYERS = 6;
constexpr int RC_PARAMS = 18624;
constexpr int RC_PADDED = 20480;
constexpr int RC_CHUNKS = 32;    // gradient partials are summed in at most this many chunks first

__host__ __device__ constexpr int rc_in(int l) { return l == 0 ? RC_IN : RC_HID; }
__host__ __device__ constexpr int rc_out(int l) { return l == RC_LAYERS - 1 ? RC_OUT : RC_HID; }
__host__ __device__ constexpr int rc_out_pad(int l) { return l == RC_LAYERS - 1 ? 32 : RC_HID; }
__host__ __device__ constexpr int rc_off(int l) {
    return l == 0 ? 0 : (l == 1 ? 2048 : (l == 2 ? 6144 : (l == 3 ? 10240 : (l == 4 ? 14336 : 18432))));
}

// Position normalisation and the radiance unit of the network output.
struct CacheEncode {
    float bmin[3];
    float binv[3];     // 1 / (bmax - bmin) per axis
    float scale;       // radiance = scale * max(output, 0)
};

// Device buffers of one context's cache.
struct CacheBuffers {
    float* w;                 // FP32 master weights [RC_PARAMS]
    float* adam_m;            // [RC_PARAMS]
    float* adam_v;            // [RC_PARAMS]
    float* grad;              // gradient of the last step [RC_PARAMS]
    uint16_t* wb;             // BF16 [RC_PADDED], row-major [out_pad][in] per layer
    uint16_t* wbt;            // BF16 [RC_PADDED], row-major [in][out_pad] per layer
    float* partial;           // per-workgroup gradient partials [max_wg][RC_PADDED]
    float* partial_sum;       // sums of consecutive chunks of them [RC_CHUNKS][RC_PADDED]
    float* loss_partial;      // per-workgroup loss sums [max_wg]
    float* stats;             // [0] loss of the last step, [1] samples of the last step
    int32_t max_wg;
    // per-frame records (filled by the wavefront shade stage)
    float4* q_pos;            // position xyz, roughness
    float4* q_dir;            // outgoing direction xyz, bit-cast path index
    float4* q_beta;           // path throughput x reflectance proxy (0.04 + 0.96 albedo)
    float4* t_pos;            // training records: position, roughness
    float4* t_dir;            // direction, bit-cast path index
    float4* t_beta;           // throughput x reflectance proxy at the vertex
    float4* t_lbefore;        // radiance of the path up to the vertex's next-event estimate
    float4* t_target;         // suffix radiance / throughput / scale
    int32_t* counters;        // [0] queries, [1] training records, [2] this frame's training
                              // threshold (of 2^24), [3] unused
    int64_t query_cap, train_cap;
};

// Per-frame switches of the wavefront stages.
struct CacheFrame {
    int32_t enabled;          // terminate non-training paths into the cache after bounce 1
    int32_t train;            // record training suffixes
    uint32_t train_threshold; // first-frame threshold: a path trains when its hash < it (of 2^24)
    int32_t adapt;            // set the threshold from the previous frame's counts
    int32_t train_samples;    // training records wanted per frame
    uint32_t frame;
    int32_t nee;              // trace the vertex's next-event estimate; the network learns the rest
    float wide_roughness;     // lobes at least this rough (or diffuse samples) lead to a cache vertex
};

}  // namespace kpt
