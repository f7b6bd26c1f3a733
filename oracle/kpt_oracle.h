/* kpt_oracle.h — C interface of the CPU oracle (test infrastructure only).
 *
 * The oracle restates the reference's host builders and device routines as
 * plain single-threaded-per-pixel C++ in IEEE semantics.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
 * reference results are compared with; the product never links it.
 */
#ifndef KPT_ORACLE_H
#define KPT_ORACLE_H
#include "../include/kpt_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

orc_scene* orc_scene_create(const kpt_scene_desc* desc);
void orc_scene_destroy(orc_scene* s);
/* [0]=triangles [1]=bvh nodes (incl. unused tail) [2]=tlas nodes [3]=instances
 * [4]=max blas depth [5]=max tlas depth [6]=lights [7]=materials */
void orc_scene_info(const orc_scene* s, int64_t out[8]);
/* BLAS of geometry mesh m in the reference layout (32-B BVH2Node), node
 * indices relative to the BLAS root, and its tri_index slice. */
int64_t orc_get_blas(const orc_scene* s, int32_t mesh, void* nodes32, int64_t cap,
                     int32_t* tri_index, int64_t cap_tris);
int64_t orc_get_tlas(const orc_scene* s, void* nodes32, int64_t cap);

/* Traversal of host rays.  mode 0 = BVH closest hit (TLAS::intersect),
 * 1 = BVH any hit (TLAS::intersectP), 2 = brute-force closest
 * (INTERSECT_DEBUG, integrator.cuh:64-95), 3 = brute-force any hit,
 * 4 = BVH closest hit entering each BLAS with the closest distance so far as
 * its tmax (the GPU kernels' bound; compared with mode 0 by a test). */
void orc_trace(const orc_scene* s, const kpt_ray* rays, int64_t n, int32_t mode,
               kpt_hit* hits, int32_t* occluded, kpt_trace_counters* counters,
               int32_t* max_stack, int32_t threads);

/* Per-pixel megakernel semantics (kernels.cu:385-451): renders samples
 * [frame0, frame0+spp) for the listed pixels with static accumulation.
 * double_precision selects the FP64 instantiation (libm transcendentals).
 * out_rgb[3*i] is the estimate after the last sample (sum / count).
 * out_first: optional [n][4] float: last sample's gbuffer bits
 * (enc nx, enc ny, albedo u32, depth) and out_v: [n][4] vbuffer. */
void orc_render_pixels(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam,
                       int32_t width, int32_t height, const int32_t* pixels, int64_t n,
                       int32_t frame0, int32_t spp, int32_t double_precision,
                       float* out_rgb, uint32_t* out_gbuffer, float* out_vbuffer,
                       int32_t threads);

/* One frame whose paths all end in the radiance cache at the bounce-1 vertex:
 * out_L [W*H][3] radiance before the cache takes over (including that
 * vertex's next-event estimate when nee != 0; unclamped), out_q
 * [W*H][12] (hit, position xyz, roughness, direction xyz, throughput xyz, 0). */
void orc_render_cache_queries(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam,
                              int32_t width, int32_t height, int32_t frame, int32_t nee, float wide_roughness,
                              float* out_L, float* out_q, int32_t threads);
/* Renderer::setMeshTransform for instance i (previous inverse kept) and the
 * TLAS rebuild. */
void orc_set_instance_transform(orc_scene* s, int32_t instance, const float model[16]);
/* One temporal-accumulation frame over the whole image: render [W*H][4]
 * (reprojected colour, history length), gbuffer [W*H][4], vbuffer [W*H][4];
 * prev_accum / prev_gbuffer are the previous frame's buffers. */
void orc_render_temporal(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam,
                         int32_t width, int32_t height, int32_t frame, const float* prev_accum,
                         const uint32_t* prev_gbuffer, float* out_render, uint32_t* out_gbuffer,
                         float* out_vbuffer, int32_t threads);

/* The same for the listed pixels only (out_* are [n][4]). */
void orc_render_temporal_pixels(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam,
                                int32_t width, int32_t height, int32_t frame, const float* prev_accum,
                                const uint32_t* prev_gbuffer, const int32_t* pixels, int64_t n,
                                float* out_render, uint32_t* out_gbuffer, float* out_vbuffer, int32_t threads);
/* FP32 op counts per phase and kind of the listed pixels' samples of frame
 * `frame` (the oracle instantiated with a counting float); out[7 * 4]. */
void orc_count_ops(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam, int32_t W, int32_t H,
                   const int32_t* pixels, int64_t n, int32_t frame, uint64_t* out);
/* Per-pixel traversal counts of one sample's closest-hit traces (DebugData,
 * blas.cuh:39-42): out[2i] BLAS node pops, out[2i+1] TLAS node pops;
 * entry_bound: BLASes entered as the GPU kernels enter them. */
void orc_heat_map(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam, int32_t W, int32_t H,
                  const int32_t* pixels, int64_t n, int32_t frame, int32_t entry_bound, int64_t* out);

/* One sample (radiance before clamp/accumulation) per pixel for frame `frame`.
 * Used to check the accumulation kernels independently. */
void orc_sample_pixels(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam,
                       int32_t width, int32_t height, const int32_t* pixels, int64_t n,
                       int32_t frame, float* out_rgb, int32_t threads);

/* BSDF probes (UnifiedBSDF, bsdf.cuh:166-624) in the shading frame
 * (basis = identity, normal +z).  params: [albedo r g b, metallic, roughness,
 * transmission, ior, backface].  dp selects the double instantiation.
 * f: out_f[3*i]; pdf: out_pdf[i]. */
void orc_bsdf_eval(const float* params, const float* wo, const float* wi, int64_t n, int32_t dp,
                   float* out_f, float* out_pdf);
/* sampleF with uniform numbers u[4*i..] = (u.x, u.y, X.x, X.y); outputs wi,
 * f, pdf and the scatter flags. */
void orc_bsdf_sample(const float* params, const float* wo, const float* u, int64_t n, int32_t dp,
                     float* out_wi, float* out_f, float* out_pdf, int32_t* out_flags);
/* GGX::D for a given alpha and microfacet normal; fresnelDielectric. */
void orc_ggx_d(float alpha, const float* wm, int64_t n, int32_t dp, float* out);
void orc_fresnel_dielectric(const float* cos_theta, const float* ior, int64_t n, int32_t dp, float* out);
/* NishitaAtmosphereModel::sampleLi for rays (o, d) with sun direction sun. */
void orc_sky_li(const float* o, const float* d, int64_t n, const float sun[3], float emission, int32_t dp,
                float* out_rgb);
/* Camera::generateRay for pixels (x, y). */
void orc_camera_rays(const kpt_camera* cam, int32_t width, int32_t height, const int32_t* pixels, int64_t n,
                     float* out_org, float* out_dir);

/* Deterministic math exported for tests. fn: 0 exp 1 exp2 2 log2 3 pow 4 sin
 * 5 cos 6 tan 7 acos 8 log */
void orc_dmath(int32_t fn, const float* x, const float* y, float* out, int64_t n);
/* pcg_hash (samplers.cuh:12-17). */
uint32_t orc_pcg_hash(uint32_t v);
/* glm::inverse (vendored glm 1.0.0 compute_inverse<4,4>) and the reference's
 * Matrix4x4::inverse (matrix_maths.cuh:304-354). */
void orc_glm_inverse(const float m[16], float out[16]);
void orc_mat4_inverse(const float m[16], float out[16]);

/* ---- post chain (kpt_oracle_post.cpp) ---- */
/* One frame of executeRendering after the path trace (renderer.cu:122-170):
 * estimate [W*H][4] (demodulated radiance) and gbuffer [W*H][4] in; modulated
 * render, post-processed output and (bloom on) bloom mip 0 out, each
 * [W*H][4] or NULL; exposure and film are updated in place when auto
 * exposure is on; out_hist (optional) receives the 256 fixed-point bins. */
void orc_post_frame(const float* estimate, const uint32_t* gbuffer, int32_t width, int32_t height,
                    const kpt_settings* settings, float frame_delta_ms, kpt_exposure* exposure,
                    kpt_film_state* film, float* out_render, float* out_output, float* out_bloom,
                    uint64_t* out_hist);
void orc_set_exposure(const kpt_exposure* e, kpt_film_state* film);
void orc_exposure_program(kpt_exposure* e, float avg_lum);
void orc_lum_to_bin(const float* lum, int64_t n, float black_ev, float white_ev, uint32_t* out);
void orc_meter_weight(int32_t width, int32_t height, const int32_t* xy, int64_t n, float* out);
void orc_histogram(const float* rgba, int32_t width, int32_t height, float white_ev, float black_ev,
                   uint64_t* bins);
float orc_histogram_average(const uint64_t* bins, float white_ev, float black_ev, float last, float dt_ms);
void orc_film(const float* rgb, int64_t n, float scalar, float white_ev, float black_ev, float* out);
void orc_post_effects(const float* back, int32_t width, int32_t height, float* out);
void orc_post_frame_sized(const float* estimate, const uint32_t* gbuffer, int32_t W, int32_t H, int32_t OW,
                          int32_t OH, int32_t upscale, const kpt_settings* s, float frame_delta_ms,
                          kpt_exposure* exposure, kpt_film_state* film, float* out_render, float* out_output,
                          float* out_bloom, uint64_t* out_hist);
void orc_upscale_bilinear(const float* src, int32_t W, int32_t H, int32_t OW, int32_t OH, float* out);
void orc_agx_curve(const float* x, int64_t n, float* out);
void orc_srgb_encode(const float* x, int64_t n, float* out);
void orc_yxy_roundtrip(const float* rgb, int64_t n, float scalar, float* out_yxy, float* out_rgb);
void orc_bilinear(const float* rgba, int32_t width, int32_t height, const float* xy, int64_t n,
                  int32_t filter_alpha, float* out);
void orc_bloom_down(const float* src, int32_t sw, int32_t sh, float* dst, int32_t dw, int32_t dh, int32_t karis);
void orc_bloom_up(const float* src, int32_t sw, int32_t sh, float* dst, int32_t dw, int32_t dh, float blend);
int32_t orc_bloom_levels(int32_t width, int32_t height);

/* ---- radiance-cache network (kpt_oracle_cache.cpp) ---- */
/* enc = {bmin xyz, binv xyz, scale}; pos/dir n x 4 floats; out n x 32 */
void orc_cache_encode(const float* pos, const float* dir, int64_t n, const float enc[7], float* out);
/* mode 0: BF16 rounding points of the GPU, double sums; mode 1: exact FP64.
 * w: KPT_CACHE_PARAMS floats; out n x 3. */
void orc_cache_forward(const float* w, const float* pos, const float* dir, int64_t n, const float enc[7],
                       int32_t mode, double* out);
/* Mean relative-L2 loss over n samples (targets n x 4) and, when grad is not
 * NULL, its gradient; w_den (or NULL) gives the denominator outputs. */
double orc_cache_loss_grad(const float* w, const float* w_den, const float* pos, const float* dir,
                           const float* targets, int64_t n, const float enc[7], int32_t mode, double* grad);
/* One Adam step (beta 0.9 / 0.99, eps 1e-15) in float, in place. */
void orc_cache_adam(float* w, float* m, float* v, const float* grad, int64_t step, float lr);

#ifdef __cplusplus
}
#endif
#endif
