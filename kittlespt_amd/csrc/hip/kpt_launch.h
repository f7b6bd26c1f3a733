// kpt_launch.h — host-side launch wrappers of the gfx950 kernels (defined in
// kpt_kernels.hip, called by the C-ABI implementation).
#pragma once
#include <vector>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/kpt_abi.h"
#include "../common/kpt_device_types.h"
#include "kpt_cache.h"

namespace kpt {

// Frame buffers of one context, device pointers.
struct FrameBuffers {
    float4* accum;     // running sum (static) or history (temporal)
    float4* render;    // current estimate
    uint4* gbuffer;
    uint4* prev_gbuffer;
    float4* vbuffer;
    float4* output;
};

// W.counters: WF_COUNTER_SLOTS queue counters, WF_COUNTER_STRIDE words
// (256 bytes) apart
constexpr int WF_COUNTER_SLOTS = 128;
constexpr int WF_COUNTER_STRIDE = 64;
constexpr int WF_COUNTER_WORDS = WF_COUNTER_SLOTS * WF_COUNTER_STRIDE;
// W.tpool: bytes per wave (64 parked traces of 256 bytes)
constexpr int TAIL_POOL_BYTES = 64 * 256;

// Per-path state of the wavefront tracer (struct of arrays in HBM).
struct WavefrontState {
    int64_t capacity;
    // path records
    float4* ray_o;        // xyz origin, w = bit-cast pixel index
    float4* ray_d;        // xyz direction, w = bit-cast seed
    float4* beta;         // xyz throughput, w = p_b
    float4* L;            // xyz radiance, w = bit-cast (depth | flags<<16)
    float4* prev_pos;     // xyz previous hit position
    float4* vis_a;        // visible surface: albedo xyz, depth
    float4* vis_n;        // normal xyz, bit-cast instance
    float4* vis_p;        // position xyz, bit-cast primitive
    // hits of the current bounce
    float4* hit;          // t, bit-cast slot, bit-cast instance, u ; v kept in hit_v
    float* hit_v;
    // queues (path indices)
    int32_t* queue[2];
    int32_t* shadow_queue;
    float4* shadow_o;     // xyz origin, w = tmax
    float4* shadow_d;     // xyz direction
    float4* shadow_c;     // contribution when unoccluded
    float4* shadow_z;     // contribution when occluded ((0*beta)/p)
    int32_t* sort_tmp;
    int32_t* counters;    // queue sizes and pull counters (WF_COUNTER_WORDS words)
    int32_t* hist;        // material histogram for the shading sort
    // spatial ray sort: keys and path/shadow indices (in, out), sort workspace
    uint32_t* rkeys[2];
    int32_t* rvals[2];
    void* rtemp;
    size_t rtemp_bytes;
    // the tail's parked closest-hit traces: TAIL_POOL slots per wave for the
    // first tpool_waves waves of its grid (null: nothing is parked)
    float4* tpool;
    int64_t tpool_waves;
    // counting mode, one view: per pixel, the BLAS and TLAS node pops of the
    // frame's closest-hit traces (DebugData, blas.cuh:39-42); null otherwise
    int32_t* heat;
};

// Exposure state resident in device memory: the auto-exposure program's
// values, the film the post pass reads and the adapted scene luminance.  The
// histogram kernels update it every frame without a host round trip.
struct ExposureState {
    kpt_exposure ae;          // AutoExposureProgram::m_exposure_values + ev comp, white/black
    float scalar;             // Film::luminance_exposure_scalar
    float white_ev;           // Film::white_point_ev
    float black_ev;           // Film::black_point_ev
    float avg_lum;            // ShaderData::scene_average_luminance
};
constexpr int MAX_MIPS = 7;   // MipChain::max_mip_count

// Buffers of the post chain (bloom mips at output resolution, histogram bins).
struct PostBuffers {
    float4* mip[MAX_MIPS];
    int32_t levels;                 // mips beyond mip 0
    unsigned long long* hist;       // 256 fixed-point bins (weight * 2^23)
    ExposureState* expo;
    float4* back;                   // back buffer when the effects pass is on (output size)
    float4* up;                     // the render buffer upscaled to the output size
};
struct PostParams {
    int32_t width, height;
    int32_t temporal;               // copy render -> accumulation before modulation
    int32_t auto_exposure;
    int32_t bloom, karis;
    float bloom_final_blend, bloom_internal_blend;
    float frame_delta_ms;
    int32_t effects;                // fisheye + chromatic aberration from pb.back into fb.output
    int32_t out_width, out_height;  // output size (width x height is the render size)
    int32_t upscale;                // bilinear upscale render -> pb.up, which bloom and post read
    int32_t fsr;                    // with upscale: FSR 1.0 (EASU render -> pb.back, RCAS -> pb.up)
};
// executeRendering after the path trace (renderer.cu:117-170): modulate,
// auto exposure, bloom, post.  Everything stays on the stream.
hipError_t launch_post_frame(hipStream_t s, const FrameBuffers& fb, const PostBuffers& pb, const PostParams& p,
                             int num_cus);
// Mip levels of the bloom chain (mipchain.cuh:17-25, 44-50).
int bloom_levels(int width, int height);

// TLAS rebuild on the device (kpt_tlas.hip): the host builder's tree for
// n <= TLAS_DEVICE_MAX instance boxes (6 floats each, finite, |v| <= 1e18), in
// the reference numbering (ref_nodes, 2n entries), the pair layout
// (pair_nodes, 2n) and the node blocks (n + 1).  status[0] = nodes built.
constexpr int TLAS_DEVICE_MAX = 2048;
hipError_t launch_tlas_build(hipStream_t s, const float* bounds, int32_t n, Node* ref_nodes, Node* pair_nodes,
                             NodeBlock* blocks, int32_t* status);
// TLAS refit on the device: the topology of ref_nodes (the last build, whose
// parent[] maps every node to its parent, 0 for the root), boxes recomputed
// from the instance boxes; then the pair layout and blocks as above.
hipError_t launch_tlas_refit(hipStream_t s, const float* bounds, int32_t n, Node* ref_nodes, const int32_t* parent,
                             int32_t* arrivals, Node* pair_nodes, NodeBlock* blocks, int32_t* status);

// Box keys hold the slot position 3*slot+vertex+1 in 32 bits and node ids
// are int32 (up to 2T+M): 2^29 triangles keeps both in range (and the
// build's scratch, ~430 bytes per triangle, inside one device's HBM).
constexpr int32_t BLAS_DEVICE_MAX_TRIS = 1 << 29;
// BLASes of all meshes built on the device, node for node the host
// builder's trees (kpt_blas.hip).  nodes: sum over meshes of 2*count-1
// entries (mesh m's slice first, unused tail zeroed, local indices as
// BlasTree); order: slot -> global triangle for all T slots; used and
// max_depth per mesh.  Blocks the host until the trees are read back.
hipError_t device_build_blas(hipStream_t st, const float* pos_host, const int32_t* mesh_first, const int32_t* mesh_count,
                             int32_t M, int32_t T, std::vector<Node>* nodes, std::vector<int32_t>* order,
                             std::vector<int32_t>* used, std::vector<int32_t>* max_depth);

hipError_t launch_trace_closest(hipStream_t s, const SceneView& S, const kpt_ray* rays, kpt_hit* hits, int64_t n,
                                int32_t* overflow, kpt_trace_counters* counters);
hipError_t launch_trace_any(hipStream_t s, const SceneView& S, const kpt_ray* rays, int32_t* occluded, int64_t n,
                            int32_t* overflow);
hipError_t launch_megakernel(hipStream_t s, const SceneView& S, const FrameConst& fc, int32_t frame,
                             const FrameBuffers& fb, int32_t* overflow);
hipError_t launch_clear(hipStream_t s, float4* p, int64_t n);

// Launch-shape knobs of the trace stages (chosen by the autotuner).
struct TraceTuning {
    int refill;        // 0: a lane traces one ray to completion; 1: idle lanes are refilled
    int chunk;         // queue entries claimed per wave atomic (refill mode)
    int refill_min;    // idle lanes that trigger a refill
    int blocks_per_cu; // persistent grid size per compute unit
    int ray_sort;      // sort secondary and shadow rays by direction octant + Morton cell of the origin
    int tail;          // from this bounce on, live paths finish in one path-per-lane launch (0: never)
    int stack;         // traversal stack entries per lane held in LDS (LDS_STACK_MIN..LDS_STACK_MAX)
    int block;         // lanes per workgroup of the persistent trace, shadow and tail kernels (64, 128, 256)
    int suspend;       // tail: park a vertex's closest-hit traces still running when fewer than this many
                       // lanes of the wave trace (0: never)
};

// Events bracketing every closest-hit trace launch of a frame (recorded on the
// context stream), and optional traversal event counters.
constexpr int WF_MAX_TRACE_LAUNCHES = 40;
// Further timed stages of a frame: the primary vertex's shade launch, the
// sky pass of the tail's escaped paths, the radiance cache's inference and
// its training steps (train, partial sums, Adam).
enum { WF_X_SHADE0 = 0, WF_X_SKY = 1, WF_X_INFER = 2, WF_X_TRAIN = 3, WF_XSTAGES = 4 };
struct WfInstrument {
    hipEvent_t begin[WF_MAX_TRACE_LAUNCHES];
    hipEvent_t end[WF_MAX_TRACE_LAUNCHES];
    int launches;                          // filled by wavefront_frame
    hipEvent_t tail_begin, tail_end;       // around the path-per-lane tail launch
    int tail_launches;                     // 0 or 1, filled by wavefront_frame
    hipEvent_t xbegin[WF_XSTAGES], xend[WF_XSTAGES];
    int xlaunches[WF_XSTAGES];             // 0 or 1 per frame
    unsigned long long* counters;          // device [16] or nullptr (counted traversal):
                                           // [0..4] closest-hit stage events, [8..12] tail
                                           // events (closest and any hit), [13] tail paths,
                                           // [14] tail vertices
};

// One camera of a multi-view batch (device array).
struct alignas(16) ViewConst {
    float inv_view[16];
    float inv_proj[16];
    float cam_pos[3];
    int32_t width;
    int32_t height;
    int32_t pad;
    int64_t offset;      // first path (pixel) of this view in the batch
};
// Many views of uneven resolution traced as one set of paths.
struct ViewBatch {
    const ViewConst* views;
    int32_t nviews;
    int64_t total;       // pixels of all views
    float4* out;         // running sum per pixel, views back to back
};
hipError_t launch_views_mean(hipStream_t s, float4* out, int64_t n, int32_t spp);
// Workspace bytes of the spatial ray sort over n rays.
size_t ray_sort_temp_bytes(int64_t n);

// The radiance cache of the wavefront tracer (null when unused).
struct CacheHooks {
    const CacheBuffers* cb;
    CacheFrame cf;
    CacheEncode enc;
    int num_cus;
};

// Wavefront path tracer: one reference frame (one sample per pixel).  With
// a cache, non-training paths end in it at the bounce-1 vertex and training
// paths leave suffix targets for the caller's training steps.
// DebugData heat map colour (kernels.cu:442-443): rgba[i] = (0, 0.02 *
// blas pops, 0.05 * tlas pops, 1) of heat[2i], heat[2i + 1].
hipError_t launch_heat_colour(hipStream_t s, const int32_t* heat, int64_t n, float4* rgba);

// finish_after (optional): an event the frame's finishing launch (the
// accumulation and G-buffer writes) waits for, so that frames whose path
// stages run on two streams still finish in frame order.  A frame with the
// radiance cache then also waits for it before its first launch that reads
// or writes the cache (its bounce-1 work: the cache vertex is never the
// primary one), and starts its cache frame there, so that its records,
// queries and inference follow the previous frame's training.
hipError_t wavefront_frame(hipStream_t s, const SceneView& S, const FrameConst& fc, int32_t frame,
                           const FrameBuffers& fb, WavefrontState& W, int32_t* overflow, int32_t num_materials,
                           int32_t sort_by_material, const TraceTuning& tune, WfInstrument* ins,
                           const CacheHooks* cache, const ViewBatch* views = nullptr,
                           hipEvent_t finish_after = nullptr);

// Radiance-cache kernels (kpt_cache.hip).
int cache_train_workgroups(int32_t batch);
hipError_t launch_cache_publish(hipStream_t s, const CacheBuffers& cb);
hipError_t launch_cache_frame_start(hipStream_t s, const CacheBuffers& cb, const CacheFrame& cf);
hipError_t launch_cache_infer_paths(hipStream_t s, const CacheBuffers& cb, const CacheEncode& enc, float4* L,
                                    int64_t max_queries, int num_cus);
hipError_t launch_cache_infer_raw(hipStream_t s, const CacheBuffers& cb, const CacheEncode& enc, const float4* pos,
                                  const float4* dir, int64_t n, float4* out, int num_cus);
hipError_t launch_cache_targets(hipStream_t s, const CacheBuffers& cb, const float4* L, float scale);
hipError_t launch_cache_train_step(hipStream_t s, const CacheBuffers& cb, const CacheEncode& enc, const float4* pos,
                                   const float4* dir, const float4* target, const int32_t* count_dev,
                                   int64_t count_host, int64_t cap, int64_t first, int32_t batch, float lr,
                                   int64_t step);

}  // namespace kpt
