/* kpt_abi.h — C-ABI of libkpt_hip.so, the MI355X (gfx950) kernel library
 * behind the KittlesPT renderer API.
 *
 * The reference has no C-ABI: its boundary is the C++ class
 * KittlesPT::Renderer (KittlesPT/src/renderer.hpp:13-58) over POD scene types
 * (KittlesPT/src/pod_types.hpp:9-219), and below it the host->kernel seam
 * launch*Kernel(const ShaderData&) (KittlesPT/src/shaders/kernels.cuh:10-19).
 * Every entry point below names the reference call it replaces.  Plain
 * pointers and sizes only; every call returns an int status (KPT_OK == 0) and
 * never exits the process (the reference calls exit(99) on any CUDA error,
 * KittlesPT/src/error_check.cu:4-13).  A context owns one HIP stream and all of
 * its device memory; it is not thread-safe, separate contexts may be used from
 * separate threads.
 */
#ifndef KPT_ABI_H
#define KPT_ABI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KPT_ABI_VERSION 1

enum kpt_status {
    KPT_OK = 0,
    KPT_ERR_INVALID_ARG = 1,
    KPT_ERR_HIP = 2,
    KPT_ERR_OUT_OF_MEMORY = 3,
    KPT_ERR_STATE = 4,        /* call order violated (e.g. render before upload) */
    KPT_ERR_OVERFLOW = 5,     /* traversal stack or queue capacity exceeded */
    KPT_ERR_NO_DEVICE = 6,
    KPT_ERR_UNSUPPORTED = 7
};

typedef struct kpt_ctx kpt_ctx;

/* MaterialSceneEntity (pod_types.hpp:60-123) without the name string. */
typedef struct kpt_material {
    int32_t albedo_texture_id;
    float albedo[3];
    int32_t orm_texture_id;
    float metallic;
    float roughness;
    int32_t transmission_texture_id;
    float transmission;
    float ior;
    int32_t emission_texture_id;
    float emission_factor[3];
    float emission_scale_nits;
    int32_t normal_texture_id;
    float normal_scale;
} kpt_material;

/* TextureSceneEntity (pod_types.hpp:163-175): 8-bit texels at `offset` bytes
 * into kpt_scene_desc.pixels. */
typedef struct kpt_texture_desc {
    uint32_t width;
    uint32_t height;
    uint32_t channels;
    uint32_t reserved;
    uint64_t offset;
} kpt_texture_desc;

/* Flattened BasicScene (pod_types.hpp:192-219).  Triangles of all meshes are
 * stored back to back; mesh m owns [mesh_first[m], mesh_first[m]+mesh_count[m]).
 * Instances place a mesh with a model matrix (column-major, glm layout).  The
 * reference has one instance per mesh and requires setMeshTransform(i, M_i)
 * after loadScene (renderer.cu:250-269, blas.cuh:190-200); the upload performs
 * both steps.  num_instances == 0 means "one instance per mesh" using
 * mesh_model. */
typedef struct kpt_scene_desc {
    int32_t num_triangles;
    const float* positions;     /* [num_triangles][3][3] */
    const float* normals;       /* [num_triangles][3][3] */
    const float* uvs;           /* [num_triangles][3][2] */
    const int32_t* material_ids;/* [num_triangles] */
    int32_t num_meshes;
    const int32_t* mesh_first;  /* [num_meshes] */
    const int32_t* mesh_count;  /* [num_meshes] */
    const float* mesh_model;    /* [num_meshes][16], used when num_instances == 0 */
    int32_t num_instances;
    const int32_t* instance_mesh;   /* [num_instances] */
    const float* instance_model;    /* [num_instances][16] */
    int32_t num_materials;
    const kpt_material* materials;
    int32_t num_textures;
    const kpt_texture_desc* textures;
    const uint8_t* pixels;
    uint64_t num_pixel_bytes;
} kpt_scene_desc;

/* A ray for the trace entry points: origin, tmin, direction, tmax. */
typedef struct kpt_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} kpt_ray;

/* Closest-hit result (Intersection, interaction.cuh:69-88). */
typedef struct kpt_hit {
    float t;            /* +inf on a miss */
    int32_t prim_id;    /* global triangle index, -1 on a miss */
    int32_t inst_id;    /* instance index, -1 on a miss */
    float bary[3];
} kpt_hit;

/* Traversal event counters; bytes follow the reference node/triangle layout. */
typedef struct kpt_trace_counters {
    uint64_t blas_interior_pops;
    uint64_t blas_leaf_pops;
    uint64_t tri_tests;
    uint64_t blas_entries;
    uint64_t tlas_pops;
    uint64_t rays;
} kpt_trace_counters;

/* Camera (camera.cuh:9-55): inverse matrices, column-major. */
typedef struct kpt_camera {
    float inv_view[16];
    float inv_proj[16];
    float prev_inv_view[16];
    float prev_inv_proj[16];
} kpt_camera;

enum kpt_seed_mode {
    KPT_SEED_REFERENCE = 0,   /* (x + y*W) * float(max(frame,1)), samplers.cuh:24-29 */
    KPT_SEED_XOR = 1          /* pixel_index ^ frame */
};

enum kpt_integrator {
    KPT_INTEGRATOR_MEGAKERNEL = 0,
    KPT_INTEGRATOR_WAVEFRONT = 1
};

/* RendererSettings (pod_types.hpp:46-57) + ProceduralEnvironmentSettings
 * (pod_types.hpp:37-44) + the extensions of this library. */
typedef struct kpt_settings {
    int32_t max_depth;
    int32_t use_temporal_accumulation;
    int32_t bloom_generate_bloom;
    int32_t bloom_use_karis_average;
    float bloom_final_blend;
    float bloom_internal_blend;
    int32_t enable_auto_exposure;
    float sun_angular_diameter_rad;
    float sun_phi_rad;
    float sun_theta_rad;
    float sun_emission_factor;
    int32_t seed_mode;        /* kpt_seed_mode */
    int32_t integrator;       /* kpt_integrator */
    /* RendererSettings::postprocess_enable_effects (pod_types.hpp:55): the
     * fisheye + chromatic-aberration pass after the post-process
     * (renderer.cu:172-178, kernels.cu:508-528); 0/1. */
    int32_t postprocess_enable_effects;
    /* RendererSettings::upscale_enable (pod_types.hpp:56): render at 1/2.5
     * of the output size and upscale bilinearly (renderer.cu:55-75, 138-160;
     * upscale.cuh:260-298 without USE_FSR); the render size
     * takes effect at the next kpt_resize (the reference's resizeResolution). */
    int32_t upscale_enable;
    /* With upscale_enable: FSR 1.0 instead of the bilinear upscale — EASU
     * from the render size to the output size, then RCAS (sharpness 0), as
     * the reference's USE_FSR build does (upscale.cuh:13-408, kernels.cu:
     * 84-103, renderer.cu:138-147; compiled out there by default,
     * kernels.cuh:3); 0/1.  The reference's inversesqrt (rsqrtf) is taken
     * as the IEEE 1/sqrtf. */
    int32_t upscale_fsr;
} kpt_settings;

/* Camera exposure as Renderer::setExposure takes it (ExposureValues,
 * pod_types.hpp:9-33, plus the EV compensation and the film's white and
 * black points, renderer.cu:312-329). */
typedef struct kpt_exposure {
    float aperture_f_num;
    float iso;
    float shutter_speed_secs;
    int32_t max_iso;
    int32_t min_iso;
    float min_shutter_secs;
    float max_shutter_secs;
    float ev_comp;
    float white_point_ev;
    float black_point_ev;
} kpt_exposure;

/* Film state the post pass reads (Film, film.cuh:186-191) and the adapted
 * scene luminance of the auto-exposure loop (histogram.cu:147-151). */
typedef struct kpt_film_state {
    float luminance_exposure_scalar;
    float white_point_ev;
    float black_point_ev;
    float scene_average_luminance;
} kpt_film_state;

/* Radiance cache (no reference counterpart: the learned part of this
 * renderer).  A 6-layer, 64-wide ReLU MLP over a 32-feature frequency
 * encoding of (position, direction, roughness); BF16 weights on the matrix
 * cores, FP32 accumulation; trained online with Adam (beta 0.9 / 0.99,
 * eps 1e-15) on a relative L2 loss over the tracer's own path suffixes. */
typedef struct kpt_cache_settings {
    int32_t enabled;          /* non-training paths end in the cache at the bounce-1 vertex */
    int32_t train;            /* record suffixes and train every frame */
    int32_t train_samples;    /* training paths per frame (65536) */
    int32_t train_steps;      /* Adam steps per frame; batch = train_samples / train_steps */
    float learning_rate;      /* 1e-3 */
    float radiance_scale;     /* radiance = radiance_scale * max(output, 0) */
    uint32_t seed;            /* weight initialisation (He uniform from a PCG hash) */
    int32_t nee_at_vertex;    /* 0: the network predicts all radiance scattered at the
                                 vertex (its next-event estimate is not traced);
                                 1: that estimate is traced exactly and the network
                                 predicts only what the continuation adds */
    float bounds_min[3];      /* position normalisation box; an empty box means */
    float bounds_max[3];      /* the scene's TLAS bounds */
    float wide_roughness;     /* a path's cache vertex is the first one after the primary hit
                                 reached through a diffuse sample or a lobe of at least this
                                 roughness (0: always the bounce-1 vertex) */
} kpt_cache_settings;

#define KPT_CACHE_PARAMS 18624

/* One camera of a multi-view batch: the matrices Renderer::setView takes
 * (column-major) and the view's resolution. */
typedef struct kpt_view {
    float projection[16];
    float view[16];
    int32_t width;
    int32_t height;
} kpt_view;

/* ---- context ---- */
int kpt_get_abi_version(void);
/* Renderer::initialize (renderer.cu:22-46). device < 0 picks the current one. */
int kpt_ctx_create(int device, kpt_ctx** out);
/* Renderer::shutdown (renderer.cu:48-53). */
int kpt_ctx_destroy(kpt_ctx* ctx);
/* Last error text of this context (never NULL). */
const char* kpt_last_error(const kpt_ctx* ctx);

/* ---- scene ---- */
/* Renderer::loadScene + setMeshTransform per instance (renderer.cu:356-436,
 * 250-269): BLAS binned-SAH and TLAS agglomerative builds (as_builder.cuh),
 * upload of nodes, triangles, instances, materials and textures to HBM. */
int kpt_upload_scene(kpt_ctx* ctx, const kpt_scene_desc* scene);
/* Renderer::setMeshTransform (renderer.cu:250-269): updates one instance
 * (previous inverse kept for motion vectors); the TLAS is rebuilt on the host
 * before the next frame or trace (updateTLAS, renderer.cu:115). */
int kpt_set_instance_transform(kpt_ctx* ctx, int32_t instance, const float model[16]);
/* setMeshTransform for instances [first, first+count) in one call (models:
 * count x 16 floats, column-major); the TLAS is rebuilt once, before the next
 * frame or trace, as updateTLAS does every frame (renderer.cu:115).  No
 * reference counterpart for the batching itself. */
int kpt_set_instance_transforms(kpt_ctx* ctx, int32_t first, int32_t count, const float* models);
/* Renderer::setMaterial (renderer.cu:207-226). */
int kpt_set_material(kpt_ctx* ctx, int32_t index, const kpt_material* m);
/* Scene statistics: [0]=triangles [1]=blas nodes [2]=tlas nodes [3]=instances
 * [4]=max BLAS depth [5]=max TLAS depth [6]=lights [7]=materials */
int kpt_scene_info(const kpt_ctx* ctx, int64_t out[8]);
/* Copies the built BVH in the reference's layout (BVH2Node, 32 B:
 * i32 child_or_first, u32 count, f32 pmin[3], f32 pmax[3]) and the
 * triangle index permutation; for parity tests of the host builders. */
int kpt_get_blas(const kpt_ctx* ctx, int32_t mesh, void* nodes32, int64_t cap_nodes,
                 int32_t* tri_index, int64_t cap_tris, int64_t out_counts[2]);
/* TLAS nodes in the reference layout (TLASNode, 32 B: f32 pmin[3], f32 pmax[3],
 * u32 left_right_index, i32 blas_index); node 0 is the root copy. */
int kpt_get_tlas(const kpt_ctx* ctx, void* nodes32, int64_t cap_nodes, int64_t* out_count);

/* ---- traversal (TLAS::intersect / intersectP, tlas.cu:8-125) ---- */
/* rays/hits are DEVICE pointers; runs on the context stream.  A ray whose
 * traversal stack would pass 128 entries (the reference's 16-entry stacks,
 * tlas.cuh:10, blas.cuh:37, overflow there) ends early; the first
 * synchronising call after it (kpt_sync, kpt_render, kpt_render_views)
 * returns KPT_ERR_OVERFLOW once. */
int kpt_trace_closest(kpt_ctx* ctx, const kpt_ray* rays, kpt_hit* hits, int64_t n);
/* occluded[i] = 1 when any triangle is hit in (tmin, tmax). */
int kpt_trace_any(kpt_ctx* ctx, const kpt_ray* rays, int32_t* occluded, int64_t n);
/* Same traversal with event counting (slower; measurement only). */
int kpt_trace_closest_counted(kpt_ctx* ctx, const kpt_ray* rays, kpt_hit* hits, int64_t n,
                              kpt_trace_counters* host_counters);
/* kpt_trace_closest_counted without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (serving loops that must not call hipMalloc, harnesses that forbid it).
 * The counters live in a fixed 256-byte device array of the library (one
 * per GPU, held until the process ends) instead of a buffer allocated per
 * call; its size does not depend on n, so no ray count is refused.  Every
 * call on a GPU shares that array and clears it first: calls that overlap on
 * one GPU (contexts of the same device used from several threads) race.
 * Refused with KPT_ERR_STATE while instance edits wait for their TLAS
 * rebuild, which allocates (kpt_trace_closest or kpt_render performs it).
 * Synchronises the context stream; same results as kpt_trace_closest_counted. */
int kpt_trace_closest_counted_noalloc(kpt_ctx* ctx, const kpt_ray* rays, kpt_hit* hits, int64_t n,
                                      kpt_trace_counters* host_counters);

/* ---- rendering (Renderer::executeRendering, renderer.cu:103-195) ---- */
/* Renderer::setRendererSettings + setProceduralEnvironmentData
 * (renderer.cu:290-305); resets accumulation. */
int kpt_set_settings(kpt_ctx* ctx, const kpt_settings* s);
int kpt_get_settings(const kpt_ctx* ctx, kpt_settings* out);
/* Renderer::setView (renderer.cu:342-354): stores glm::inverse of both
 * matrices, the previous camera becomes prev (Camera::setView,
 * camera.cuh:28-36), resets accumulation unless temporal. */
int kpt_set_view(kpt_ctx* ctx, const float projection[16], const float view[16]);
/* Direct access to the camera inverses (current and previous). */
int kpt_set_camera(kpt_ctx* ctx, const kpt_camera* cam);
int kpt_get_camera(const kpt_ctx* ctx, kpt_camera* out);
/* Renderer::resizeResolution (renderer.cu:55-101): the output size is
 * width x height, the render size the same or (int)(width / 2.5f) x
 * (int)(height / 2.5f) when upscale_enable is set; (re)allocates the frame
 * buffers and, as the reference, recomputes the projection for the render
 * size keeping its field of view (near 0.1, far 100). */
int kpt_resize(kpt_ctx* ctx, int32_t width, int32_t height);
/* Sizes after the last kpt_resize: out[0..1] render width/height (buffers 0-3
 * of kpt_readback), out[2..3] output width/height (buffer 4).  No reference
 * counterpart (the reference keeps them in ShaderData). */
int kpt_get_resolution(const kpt_ctx* ctx, int32_t out[4]);
/* Renderer::resetAccumulation (renderer.cu:336-340). */
int kpt_reset_accumulation(kpt_ctx* ctx);
/* Renders `spp` reference frames of one sample each (static accumulation
 * semantics of computePathTraceSamplesMegaKernel, kernels.cu:385-451) starting
 * at the current frame index, then advances it by spp. */
int kpt_render(kpt_ctx* ctx, int32_t spp);
/* Frame index (ShaderData::frame_index, containers.cuh:40). */
int kpt_get_frame_index(const kpt_ctx* ctx, int32_t* out);
/* Copies a frame buffer to host memory. which: 0 = render (estimate, RGBA32F),
 * 1 = accumulation (running sum, RGBA32F), 2 = gbuffer (packed, 4xu32),
 * 3 = vbuffer (RGBA32F), 4 = output (post-processed sRGB, RGBA32F). */
int kpt_readback(kpt_ctx* ctx, int32_t which, void* host_dst, int64_t bytes);
/* Synchronises the context stream; KPT_ERR_OVERFLOW when a traversal since
 * the last such report overflowed its stack (reported once). */
int kpt_sync(kpt_ctx* ctx);
/* Milliseconds of the last kpt_render (HIP events on the context stream) and
 * of its dominant closest-hit trace kernel launches. out[0]=render ms,
 * out[1]=trace ms summed, out[2]=trace launches. */
int kpt_last_timing(const kpt_ctx* ctx, double out[4]);
/* Wall milliseconds of the last kpt_upload_scene's BLAS builds (out[0]) and
 * whether they ran on the device (out[1] = 1) or on host threads (0).  The
 * reference builds on the host inside loadScene (renderer.cu:356-436) and
 * reports no timing. */
int kpt_last_build_timing(const kpt_ctx* ctx, double out[2]);
/* Measurement mode: when enabled, the wavefront's closest-hit launches count
 * traversal events; kpt_get_render_counters returns the totals of the last
 * kpt_render.  No reference counterpart (the reference keeps per-ray
 * DebugData counters, blas.cuh:39-42). */
int kpt_set_trace_counting(kpt_ctx* ctx, int32_t enable);
/* DebugData heat map (blas.cuh:39-42 counts, kernels.cu:442-443 colours) of
 * the last frame of the last kpt_render made with trace counting on and the
 * wavefront integrator: per pixel, the BLAS node pops (blas_hits,
 * blas.cuh:140) and TLAS node pops (tlas_hits, tlas.cu:87) of its path's
 * closest-hit traces, shadow rays excluded (integrator.cuh:129).  BLASes are
 * entered with the closest hit so far as their bound, so the counts are at
 * most the reference's.  counts (or NULL): n_pixels x 2 int32; rgba (or
 * NULL): n_pixels x 4 floats, (0, 0.02 * blas pops, 0.05 * tlas pops, 1).
 * n_pixels must be that render's width x height; KPT_ERR_STATE when no
 * counted render has run.  Synchronises the context stream. */
int kpt_get_heat_map(kpt_ctx* ctx, int32_t* counts, float* rgba, int64_t n_pixels);
/* Launch-shape knobs of the wavefront stages, chosen by the autotuner (no
 * reference counterpart): knobs[0] lane refill (0/1, default 0), [1] queue
 * entries a wave claims per atomic (64..65536, default 64; without refill a
 * wave traces its claim 64 rays at a time), [2] idle lanes that trigger a refill
 * (1..64), [3] persistent workgroups per compute unit (1..64), [4] sort
 * shading by material (0 no, 1 every bounce, 2 from bounce 1 on), [5] (when n >= 6) capture every frame as a
 * hipGraph and replay it (0/1; the per-launch trace times of kpt_last_timing
 * are then not measured), [6] (when n >= 7) sort continuation and shadow
 * rays by direction octant and Morton cell of the origin before tracing
 * (0/1; not applied inside captured graphs), [7] (when n >= 8) the bounce
 * from which the paths still alive finish in one persistent launch, a lane
 * running one path at a time vertex by vertex and taking the next queued path
 * when its own ends (0: never; 0..64; default 1; on radiance-cache frames
 * the tail also finishes the cached bounce-1 vertices), [8] (when n >= 9)
 * rebuild the TLAS after instance edits on the device (1: the same tree,
 * built and laid out on the stream without a host synchronisation; slower
 * than the host builder for 1,000 instances), refit it on the device (2: the
 * topology of the last build, boxes recomputed; not the reference's tree once
 * instances move) or rebuild it on the host (0, default), [9] (when n >= 10)
 * build the BLASes of the next kpt_upload_scene on the device (1, the
 * default of a context with a device: the host builder's trees node for
 * node, all meshes in one level-by-level build) or on host threads (0, the
 * only choice of a host-only context).  A scene the device builder cannot
 * take (non-finite vertex positions, more than 2^29 triangles) or a device
 * build that fails is built on host threads; kpt_last_build_timing says
 * which ran, [10] (when n >= 11) traversal stack entries per lane held in
 * LDS by the wavefront's trace, shadow and tail kernels (8..32, default 16;
 * deeper entries go to private memory, 128 entries in all whatever the
 * split; the reference's stacks hold 16, tlas.cuh:10, blas.cuh:37), [11]
 * (when n >= 12) lanes per workgroup of those kernels (64, 128 or 256,
 * default 256; the reference launches 8x8 blocks, kernels.cu:32-36; the
 * persistent grid keeps blocks_per_cu x 256 lanes per compute unit), [12]
 * (when n >= 13) the tail's parking threshold (0..64, default 16; 0 never
 * parks): once fewer than this many lanes of a tail wave are still in a
 * vertex's closest-hit trace, and the wave still claims new paths, the
 * unfinished traces are parked (their path state in the path arrays, their
 * traversal state and stack, up to 48 entries, in a per-wave pool) and taken
 * back by the wave's next idle lanes; the image does not change, [13] (when
 * n >= 14) two frames in flight (1, default) or one (0): a render of several
 * frames runs every other frame's path stages on a second stream with its
 * own path state, each frame's accumulation and post chain after the
 * previous frame's, so the outputs are the one-at-a-time frames' (frames
 * that train or query the radiance cache, counted frames and captured graphs
 * run one at a time).
 * n >= 5. */
int kpt_set_tuning(kpt_ctx* ctx, const int32_t* knobs, int32_t n);
int kpt_get_tuning(const kpt_ctx* ctx, int32_t* knobs, int32_t n);
int kpt_get_render_counters(const kpt_ctx* ctx, kpt_trace_counters* out);
/* Per-stage measurements of the last wavefront kpt_render (no reference
 * counterpart): stage 0 = the closest-hit trace launches, stage 1 = the
 * path-per-lane tail launches, 2 = the primary vertex's shade launch, 3 = the
 * sky pass of the tail's escaped paths, 4 = the radiance cache's inference,
 * 5 = its training steps (MLP forward/backward, partial sums, Adam).  ms and launches come from HIP events around
 * each launch (zero when the frames were replayed as hipGraphs); events,
 * paths and vertices are filled in counting mode (kpt_set_trace_counting):
 * the traversal events of the stage (the tail's closest and any-hit
 * traversals together) and, for the tail, the paths it finished and the
 * path vertices it processed. */
typedef struct kpt_stage_stats {
    double ms;
    int32_t launches;
    int32_t reserved;
    kpt_trace_counters events;
    uint64_t paths;
    uint64_t vertices;
} kpt_stage_stats;
int kpt_get_stage_stats(const kpt_ctx* ctx, int32_t stage, kpt_stage_stats* out);

/* ---- post chain (executeRendering after the path trace, renderer.cu:117-186) ---- */
/* Renderer::setExposure (renderer.cu:312-329): records the values for the
 * auto-exposure program and sets the film's exposure scalar
 * (getStandardOutputBasedExposure, autoexposure_program.hpp:38-47) and
 * white/black points.  Does not reset accumulation. */
int kpt_set_exposure(kpt_ctx* ctx, const kpt_exposure* e);
/* Renderer::getExposure (renderer.cu:331-334): the auto-exposure program's
 * current values (updated on the device every frame when auto exposure is
 * on; this call synchronises the context stream). */
int kpt_get_exposure(kpt_ctx* ctx, kpt_exposure* out);
/* Film scalar, white/black points and the adapted scene luminance
 * (synchronises the context stream). */
int kpt_get_film(kpt_ctx* ctx, kpt_film_state* out);
/* executeRendering's delta_time_ms (renderer.cu:103-105): the frame time the
 * auto-exposure adaptation uses for the following kpt_render frames. */
int kpt_set_frame_delta(kpt_ctx* ctx, float delta_time_ms);

/* ---- batched views (no reference counterpart: the reference renders one
 * camera per executeRendering) ---- */
/* Renders `spp` samples per pixel of n views of any resolutions as ONE set of
 * paths: every wavefront stage's persistent kernels pull work from a single
 * queue over all views' pixels.  Sample k of a pixel uses frame index k and
 * the pixel's own seed, so each view's result equals rendering that view
 * alone with kpt_render from frame 0.  out: DEVICE RGBA32F, views back to
 * back (view v starts at the sum of the earlier views' width*height), the
 * static-accumulation estimate after spp samples.  No G-buffer, post chain
 * or cache; the context's frame buffers are untouched.  With the graph
 * tuning knob each sample's launches are captured and replayed as one
 * hipGraph (knobs as for kpt_render). */
int kpt_render_views(kpt_ctx* ctx, const kpt_view* views, int32_t n, int32_t spp, float* out);
/* The same into HOST memory (sum of width*height RGBA32F; a temporary
 * device buffer is allocated and freed per call). */
int kpt_render_views_host(kpt_ctx* ctx, const kpt_view* views, int32_t n, int32_t spp, float* host_out);

/* ---- radiance cache (no reference counterpart) ---- */
/* Applies the settings; the first call (and any call with a new seed)
 * initialises the weights and clears the Adam state.  The cache runs in the
 * wavefront integrator only. */
int kpt_cache_configure(kpt_ctx* ctx, const kpt_cache_settings* s);
int kpt_cache_get_settings(const kpt_ctx* ctx, kpt_cache_settings* out);
/* FP32 master weights, KPT_CACHE_PARAMS floats: per layer row-major
 * [out][in], layers 32->64, 4 x 64->64, 64->3.  Setting them clears Adam. */
int kpt_cache_get_weights(kpt_ctx* ctx, float* host_out, int64_t n);
int kpt_cache_set_weights(kpt_ctx* ctx, const float* host_w, int64_t n);
/* Network outputs for n queries (DEVICE pointers): positions n x 4 floats
 * (xyz, roughness), directions n x 4 (outgoing direction xyz, unused); out
 * n x 4 floats (the 3 raw outputs, 0). */
int kpt_cache_infer(kpt_ctx* ctx, const float* positions, const float* directions, int64_t n, float* out);
/* `steps` Adam steps over n samples (DEVICE pointers; queries as above,
 * targets n x 4 in output units), batches of ceil(n / steps) in order.
 * host_loss (optional) receives the loss of the last step. */
int kpt_cache_train(kpt_ctx* ctx, const float* positions, const float* directions, const float* targets,
                    int64_t n, int32_t steps, float* host_loss);
/* Gradient of the last Adam step (before the update), KPT_CACHE_PARAMS floats. */
int kpt_cache_get_gradient(kpt_ctx* ctx, float* host_out, int64_t n);
/* out[0] loss of the last step, [1] samples in it, [2] queries of the last
 * frame, [3] training records of the last frame, [4] Adam steps so far. */
int kpt_cache_stats(kpt_ctx* ctx, double out[5]);
/* Host copies of the last frame's cache queries (in slot order): positions
 * n x 4 (xyz, roughness), directions n x 4 (outgoing xyz, bit-cast path
 * index), throughputs n x 4.  *count receives the number of queries; at most
 * cap are copied. */
int kpt_cache_get_queries(kpt_ctx* ctx, float* pos, float* dir, float* beta, int64_t cap, int64_t* count);

/* ---- image data (no reference counterpart: the reference's importer
 * decodes images with stb_image, SampleApp model_importer.hpp) ---- */
/* PNG scanline unfiltering (PNG specification, section 9): raw holds height
 * rows of 1 filter byte + stride bytes (the inflated IDAT stream of an
 * 8-bit, non-interlaced image with bpp bytes per pixel); out receives
 * height x stride bytes.  KPT_ERR_INVALID_ARG on an unknown filter type. */
int kpt_png_unfilter(const uint8_t* raw, int64_t height, int64_t stride, int32_t bpp, uint8_t* out);
/* PNG image (n bytes at data) to RGB8, as the reference's importer calls
 * stbi_load_from_memory(..., 3) (SampleApp model_importer.hpp:287-296): grey
 * replicated, alpha dropped, palette looked up.  8-bit non-interlaced images
 * of colour type 0, 2, 3, 4 or 6 are read (others: KPT_ERR_UNSUPPORTED; a
 * malformed stream: KPT_ERR_INVALID_ARG).  *width and *height receive the
 * size; with rgb NULL nothing else is done, else rgb receives width x height
 * x 3 bytes, rows top-down (capacity >= that). */
int kpt_png_decode_rgb8(const uint8_t* data, int64_t n, int32_t* width, int32_t* height, uint8_t* rgb,
                        int64_t capacity);
/* ---- frame files (no reference counterpart: the reference shows frames
 * through GL textures, renderer.cu:197-205) ---- */
/* px: width x height pixels of `channels` (3 or 4) floats, row 0 = the frame
 * buffers' bottom row.  kpt_write_png_rgba8: 8-bit RGBA PNG of values in
 * [0, 1] (floor(v * 255 + 0.5) clamped, NaN -> 0, alpha 1 when absent), rows
 * flipped to top-down when flip_y.  kpt_write_pfm: colour PFM, little-endian,
 * rows as given (PFM is bottom-up).  kpt_write_exr: single-part scanline
 * OpenEXR, uncompressed FLOAT A/B/G/R, rows flipped when flip_y.
 * KPT_ERR_INVALID_ARG on bad arguments or a file that cannot be written. */
int kpt_write_png_rgba8(const char* path, const float* px, int32_t width, int32_t height, int32_t channels,
                        int32_t flip_y);
int kpt_write_pfm(const char* path, const float* px, int32_t width, int32_t height, int32_t channels);
int kpt_write_exr(const char* path, const float* px, int32_t width, int32_t height, int32_t channels,
                  int32_t flip_y);

/* ---- measurement aid (no reference counterpart) ----
 * dst[i] = src[idx[i]] for i < n: 16-byte records, one lane per record, on the
 * default stream, synchronised before return.  The traversal kernels' memory
 * pattern — one 16-byte load per lane from an address unrelated to its
 * neighbours' — with a known byte count, so that rocprofv3's FETCH_SIZE of
 * scattered reads can be calibrated against it.  src, idx and dst are device
 * pointers; an index outside [0, src_records) reads record 0.
 * KPT_ERR_INVALID_ARG on null pointers or negative sizes, KPT_ERR_HIP when
 * the launch fails. */
int kpt_measure_gather16(const void* src, int64_t src_records, const int32_t* idx, int64_t n, void* dst);

#ifdef __cplusplus
}
#endif
#endif /* KPT_ABI_H */
