// kpt_abi.hip — implementation of the C-ABI in include/kpt_abi.h: context,
// scene upload (host builders + HBM layout), render orchestration.
#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <mutex>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/kpt_abi.h"
#include "../host/kpt_build.h"
#include "../host/kpt_hostmath.h"
#include "kpt_integrator.h"
#include "kpt_launch.h"

using namespace kpt;

namespace {

template <class T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
};

}  // namespace

// Device scratch of kpt_trace_closest_counted_noalloc: the traversal event
// counters of one call (five 64-bit words used), a fixed 256 bytes whatever
// the number of rays.  Every GPU holds its own copy until the process ends;
// every call on a GPU shares it.  In host code `scratch` is only a stand-in
// for the symbol: kernels and copies take the address kept in
// g_scratch_addr, looked up once per device with that device current.
constexpr int NOALLOC_SCRATCH_WORDS = 32;
static __device__ __align__(256) unsigned long long scratch[NOALLOC_SCRATCH_WORDS];

namespace {
std::mutex g_scratch_mu;
std::vector<unsigned long long*> g_scratch_addr;   // by device; nullptr until looked up

// The device address of `scratch` on `device` (current at the call).
hipError_t scratch_address(int device, unsigned long long** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)device >= g_scratch_addr.size()) g_scratch_addr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_addr[(size_t)device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[(size_t)device] = static_cast<unsigned long long*>(ptr);
    }
    *out = g_scratch_addr[(size_t)device];
    return hipSuccess;
}
}  // namespace

struct kpt_ctx {
    int device = -1;
    bool host_only = false;
    hipStream_t stream = nullptr;
    std::string err;
    // host copy of the scene
    std::vector<float> pos, uv;
    std::vector<int32_t> mat_ids, mesh_first, mesh_count, inst_mesh;
    std::vector<float> inst_model;
    std::vector<kpt::Material> materials;
    std::vector<kpt::Texture> textures;
    std::vector<BlasTree> blas;
    std::vector<int32_t> geom_base;           // global node index of each BLAS root
    std::vector<Aabb> inst_bounds;
    TlasTree tlas;
    std::vector<Node> nodes;                  // global BLAS node array
    std::vector<InstTrav> itrav;
    std::vector<InstShade> ishade;
    std::vector<kpt::Light> lights;
    std::vector<float> light_verts;
    int64_t num_pixel_bytes = 0;
    bool scene_ready = false;
    // device scene
    Node* d_blas = nullptr;
    Node* d_tlas = nullptr;          // TLAS in the pair layout of the traversal
    NodeBlock* d_bblocks = nullptr;  // two-level records of the BLAS pairs
    NodeBlock* d_tblocks = nullptr;  // two-level records of the TLAS pairs
    TriPos* d_tris = nullptr;
    TriAttr* d_attrs = nullptr;
    InstTrav* d_itrav = nullptr;
    InstShade* d_ishade = nullptr;
    kpt::Material* d_mat = nullptr;
    kpt::Texture* d_tex = nullptr;
    uint8_t* d_pix = nullptr;
    kpt::Light* d_lights = nullptr;
    float* d_lverts = nullptr;
    int32_t* d_overflow = nullptr;
    unsigned long long* scratch_addr = nullptr;   // this device's copy of `scratch`
    // rendering
    kpt_settings settings{};
    kpt_camera cam{};
    int32_t frame_index = 0;
    int32_t width = 0, height = 0;           // render size (the output size / 2.5 when upscaling)
    int32_t out_width = 0, out_height = 0;   // output size (resizeResolution's arguments)
    FrameBuffers fb{};
    WavefrontState ws{};
    // a second frame in flight (the overlap knob): its path state and stream,
    // and per stream the event marking its last frame's finished outputs
    int overlap = 1;
    WavefrontState ws2{};
    hipStream_t stream2 = nullptr;
    hipEvent_t lane_done[2] = {nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr, tev0 = nullptr, tev1 = nullptr;
    static constexpr int INS_RING = 4;   // frames whose trace events may be in flight
    WfInstrument ins[INS_RING]{};
    int32_t use_graph = 0;               // capture each frame as a hipGraph and replay it
    hipGraphExec_t graph_exec = nullptr;        // kpt_render's per-frame graph
    hipGraphExec_t views_graph_exec = nullptr;  // kpt_render_views' per-sample graph
    TraceTuning tune{0, 64, 16, 8, 0, 1, 16, 256, 16};    // the autotuner's usual choice on the 1080p workload
    int32_t sort_by_material = 2;
    bool count_trace = false;
    unsigned long long* d_counts = nullptr;
    // DebugData heat map of the last counted render's last frame (2 int32 per pixel)
    int32_t* d_heat = nullptr;
    int64_t heat_pixels = 0;
    bool heat_valid = false;
    kpt_trace_counters render_counts{};
    double timing[4] = {0, 0, 0, 0};
    kpt_stage_stats stages[6]{};         // closest-hit trace, path-per-lane tail, primary shade, sky pass,
                                         // cache inference, cache training
    int32_t blas_device = 0;             // build the BLASes of the next upload on the device (kpt_blas.hip); 1 on device contexts
    double blas_build_ms[2] = {0, 0};    // last upload: BLAS build wall ms, 1 if built on the device
    // TLAS rebuilds on the device (kpt_tlas.hip); the host tree is then stale
    int32_t tlas_device = 0;
    bool tlas_on_device = false;
    float* d_inst_bounds = nullptr;
    Node* d_tlas_ref = nullptr;          // the device build in the reference numbering
    int32_t* d_tlas_status = nullptr;
    int32_t* d_tlas_parent = nullptr;    // refit: parent of every node of the last build (0 for the root)
    int32_t* d_tlas_arrivals = nullptr;  // refit: per-node arrival counters
    float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {1, 1, 1};   // box of all instances (ray-sort keys)
    bool tlas_dirty = false;
    bool cache_adapt = false;   // the last frame trained the cache at this resolution: its counts set the next threshold
    ViewConst* d_views = nullptr;    // multi-view batch cameras
    // post chain
    PostBuffers pb{};
    ExposureState h_expo{};
    float frame_delta_ms = 0.0f;
    int num_cus = 256;
    // radiance cache
    kpt_cache_settings cache_set{};
    CacheBuffers cb{};
    bool cache_weights = false;      // weights initialised
    int64_t cache_steps = 0;         // Adam steps taken
    double cache_frame_counts[2] = {0, 0};
};

namespace {

int fail(kpt_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(ctx, KPT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T> hipError_t dev_upload(T** dst, const std::vector<T>& src) {
    if (*dst) { (void)hipFree(*dst); *dst = nullptr; }
    size_t bytes = std::max<size_t>(sizeof(T), src.size() * sizeof(T));
    hipError_t e = hipMalloc((void**)dst, bytes);
    if (e != hipSuccess) return e;
    if (!src.empty()) e = hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
}
template <class T> void dev_free(T** p) { if (*p) { (void)hipFree(*p); *p = nullptr; } }

void default_settings(kpt_settings* s) {
    std::memset(s, 0, sizeof(*s));
    // RendererSettings / ProceduralEnvironmentSettings defaults (pod_types.hpp:37-57)
    s->max_depth = 3;
    s->bloom_final_blend = 0.175f;
    s->bloom_internal_blend = 0.65f;
    s->sun_angular_diameter_rad = 0.0087f;
    s->sun_phi_rad = 4.18879f;
    s->sun_theta_rad = 17.0f * (3.14159265358979323846f / 180.0f);
    s->sun_emission_factor = 1.0f;
    s->seed_mode = KPT_SEED_REFERENCE;
    s->integrator = KPT_INTEGRATOR_MEGAKERNEL;
}

SceneView scene_view(const kpt_ctx* c) {
    SceneView S;
    S.blas_nodes = c->d_blas;
    S.tlas_nodes = c->d_tlas;
    S.blas_blocks = c->d_bblocks;
    S.tlas_blocks = c->d_tblocks;
    S.tris = c->d_tris;
    S.attrs = c->d_attrs;
    S.inst_trav = c->d_itrav;
    S.inst_shade = c->d_ishade;
    S.materials = c->d_mat;
    S.textures = c->d_tex;
    S.pixels = c->d_pix;
    S.lights = c->d_lights;
    S.light_verts = c->d_lverts;
    S.num_pixel_bytes = c->num_pixel_bytes;
    S.tlas_root = c->tlas.root;
    S.num_lights = (int32_t)c->lights.size();
    return S;
}

void store_mat(const hm::Mat4& m, float* dst) { std::memcpy(dst, m.m, 64); }

// Two-level records of a pair-layout node array (see NodeBlock): record q
// holds nodes 2q, 2q+1 and the children of each interior one.  `interior`
// tells the level's interior test; `pair` gives an interior node's children
// pair index.  Returns false when a children pair does not sit at an even
// index (the layout the traversal relies on).
template <class IsInterior>
bool make_blocks(const std::vector<Node>& nodes, IsInterior interior, std::vector<NodeBlock>* out) {
    out->assign(nodes.size() / 2 + 1, NodeBlock{});
    for (size_t i = 0; i < nodes.size(); ++i) {
        const Node& x = nodes[i];
        if (!interior(x)) continue;
        const uint32_t p = x.a;
        if ((p & 1u) || (size_t)p + 1 >= nodes.size()) return false;
        NodeBlock& blk = (*out)[p >> 1];
        blk.c[0] = nodes[p];
        blk.c[1] = nodes[p + 1];
        for (int k = 0; k < 2; ++k) {
            const Node& ch = nodes[p + k];
            if (!interior(ch)) continue;
            if ((ch.a & 1u) || (size_t)ch.a + 1 >= nodes.size()) return false;
            blk.g[2 * k] = nodes[ch.a];
            blk.g[2 * k + 1] = nodes[ch.a + 1];
        }
    }
    return true;
}

// The TLAS renumbered into the pair layout: root at 0, index 1 unused; the
// two children of the merged node with reference index i (i > N, the
// builder's numbering: leaves 1..N, merged nodes from N + 1, node 0 a copy
// of the root, which is the last merged node 2N - 1) at nodes 2(i - N) and
// 2(i - N) + 1, first child even; an interior node's a = its children's first
// index, a leaf's a = 0 (b = instance); boxes unchanged.  Each node's place
// follows from its parent alone, so the device rebuild (kpt_tlas.hip) lays
// the tree out identically in parallel.
std::vector<Node> tlas_pair_layout(const std::vector<Node>& orig) {
    std::vector<Node> out;
    if (orig.empty()) return out;
    const uint32_t N = (uint32_t)(orig.size() / 2);
    out.assign(orig.size(), Node{});
    auto relink = [&](Node n, uint32_t self) {
        if (n.a != 0u) n.a = 2u * (self - N);
        return n;
    };
    Node root = orig[0];
    if (root.a != 0u) root.a = 2u * (N - 1u);
    out[0] = root;
    for (uint32_t i = N + 1; i < (uint32_t)orig.size(); ++i) {
        const uint32_t q = i - N, na = orig[i].a & 0xFFFFu, nb = orig[i].a >> 16;
        out[2 * q] = relink(orig[na], na);
        out[2 * q + 1] = relink(orig[nb], nb);
    }
    return out;
}

// Per-instance records from a model matrix (Renderer::setMeshTransform,
// renderer.cu:250-269; BLAS2::setTransform, blas.cuh:190-200).
void make_instance(kpt_ctx* c, int32_t i, const float* model, bool keep_prev) {
    hm::Mat4 M = hm::load(model);
    hm::Mat4 trav = hm::ref_inverse(M);
    InstTrav& it = c->itrav[i];
    for (int k = 0; k < 4; ++k) { it.r0[k] = trav.m[4 * k + 0]; it.r1[k] = trav.m[4 * k + 1]; it.r2[k] = trav.m[4 * k + 2]; }
    int32_t g = c->inst_mesh[i];
    it.root = c->blas[g].used > 0 ? c->geom_base[g] : -1;
    it.geom = g;
    it.root_entry = 0xFFFFFFFFu;
    if (it.root >= 0) {
        const Node& r = c->nodes[(size_t)it.root];
        it.root_entry = blas_entry((uint32_t)it.root, r.a, r.b);
    }
    InstShade& sh = c->ishade[i];
    hm::Mat4 curr_inv = hm::glm_inverse(M);
    hm::Mat4 prev_inv = keep_prev ? hm::load(sh.curr_inv) : curr_inv;
    store_mat(hm::ref_inverse(curr_inv), sh.model);
    store_mat(curr_inv, sh.curr_inv);
    store_mat(hm::ref_inverse(prev_inv), sh.prev_model);
    c->inst_bounds[i] = transform_bounds(c->blas[g].bounds, model);
}

void set_scene_box(kpt_ctx* ctx) {
    for (int k = 0; k < 3; ++k) { ctx->scene_lo[k] = 0.0f; ctx->scene_hi[k] = 1.0f; }
    if (ctx->inst_bounds.empty()) return;
    for (int k = 0; k < 3; ++k) { ctx->scene_lo[k] = ctx->inst_bounds[0].lo[k]; ctx->scene_hi[k] = ctx->inst_bounds[0].hi[k]; }
    for (const Aabb& b : ctx->inst_bounds)
        for (int k = 0; k < 3; ++k) {
            ctx->scene_lo[k] = std::min(ctx->scene_lo[k], b.lo[k]);
            ctx->scene_hi[k] = std::max(ctx->scene_hi[k], b.hi[k]);
        }
}

int upload_tlas(kpt_ctx* ctx) {
    ctx->tlas = build_tlas(ctx->inst_bounds);
    ctx->tlas_dirty = false;
    ctx->tlas_on_device = false;
    set_scene_box(ctx);
    if (ctx->tlas.nodes.size() > 65535) return fail(ctx, KPT_ERR_UNSUPPORTED, "TLAS exceeds 65535 nodes (16-bit child fields)");
    if (ctx->host_only) return KPT_OK;
    const std::vector<Node> tp = tlas_pair_layout(ctx->tlas.nodes);
    std::vector<NodeBlock> tb;
    if (!make_blocks(tp, [](const Node& n) { return n.a != 0u; }, &tb))
        return fail(ctx, KPT_ERR_UNSUPPORTED, "TLAS pair layout failed");
    // the previous frame may still read the old arrays
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // the device rebuild's buffers follow the instance count of this tree
    dev_free(&ctx->d_inst_bounds);
    dev_free(&ctx->d_tlas_ref);
    dev_free(&ctx->d_tlas_status);
    dev_free(&ctx->d_tlas_parent);
    dev_free(&ctx->d_tlas_arrivals);
    HIPCHK(dev_upload(&ctx->d_tlas, tp));
    HIPCHK(dev_upload(&ctx->d_tblocks, tb));
    HIPCHK(dev_upload(&ctx->d_itrav, ctx->itrav));
    HIPCHK(dev_upload(&ctx->d_ishade, ctx->ishade));
    if (ctx->tlas_device == 2 && !ctx->tlas.nodes.empty()) {
        // refit mode: the tree just built, with its parents, kept on the device
        const size_t n2 = ctx->tlas.nodes.size();
        std::vector<int32_t> parent(n2, 0);
        for (size_t i = 1; i < n2; ++i) {
            const uint32_t a = ctx->tlas.nodes[i].a;
            if (a == 0u) continue;
            parent[a & 0xFFFFu] = (int32_t)i;
            parent[a >> 16] = (int32_t)i;
        }
        HIPCHK(hipMalloc((void**)&ctx->d_inst_bounds, (n2 / 2) * 6 * sizeof(float)));
        HIPCHK(hipMalloc((void**)&ctx->d_tlas_status, sizeof(int32_t)));
        HIPCHK(hipMalloc((void**)&ctx->d_tlas_arrivals, n2 * sizeof(int32_t)));
        HIPCHK(dev_upload(&ctx->d_tlas_ref, ctx->tlas.nodes));
        HIPCHK(dev_upload(&ctx->d_tlas_parent, parent));
    }
    return KPT_OK;
}

// The TLAS rebuilt on the device, stream-ordered behind the frames that read
// the old one (no host synchronisation): instance boxes and records copied
// up, the tree built, laid out and blocked by kpt_tlas.hip.  The host
// builder runs instead for an empty or too large instance set and for boxes
// whose half areas could overflow (the device scan assumes finite ones).
int rebuild_tlas_device(kpt_ctx* ctx) {
    const int32_t N = (int32_t)ctx->inst_bounds.size();
    const bool refit = ctx->tlas_device == 2;
    bool ok = ctx->tlas_device && !ctx->host_only && N >= 1 && (refit || N <= TLAS_DEVICE_MAX) &&
              ctx->d_tlas && ctx->d_tblocks && (int64_t)ctx->tlas.nodes.size() == 2 * (int64_t)N &&
              (!refit || (ctx->d_tlas_parent && ctx->d_tlas_ref));
    for (int32_t i = 0; ok && i < N; ++i)
        for (int k = 0; k < 3; ++k) {
            const float lo = ctx->inst_bounds[i].lo[k], hi = ctx->inst_bounds[i].hi[k];
            if (!(std::fabs(lo) <= 1e18f) || !(std::fabs(hi) <= 1e18f)) ok = false;
        }
    if (!ok) return upload_tlas(ctx);
    if (refit) {
        // the topology of the last host build, boxes from the current instances
        HIPCHK(hipMemcpyAsync(ctx->d_inst_bounds, ctx->inst_bounds.data(), (size_t)N * sizeof(Aabb),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->d_itrav, ctx->itrav.data(), ctx->itrav.size() * sizeof(InstTrav),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->d_ishade, ctx->ishade.data(), ctx->ishade.size() * sizeof(InstShade),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(launch_tlas_refit(ctx->stream, ctx->d_inst_bounds, N, ctx->d_tlas_ref, ctx->d_tlas_parent,
                                 ctx->d_tlas_arrivals, ctx->d_tlas, ctx->d_tblocks, ctx->d_tlas_status));
        ctx->tlas_dirty = false;
        ctx->tlas_on_device = true;
        set_scene_box(ctx);
        return KPT_OK;
    }
    if (!ctx->d_inst_bounds) {
        HIPCHK(hipMalloc((void**)&ctx->d_inst_bounds, (size_t)N * 6 * sizeof(float)));
        HIPCHK(hipMalloc((void**)&ctx->d_tlas_ref, (size_t)2 * N * sizeof(Node)));
        HIPCHK(hipMalloc((void**)&ctx->d_tlas_status, sizeof(int32_t)));
    }
    static_assert(sizeof(Aabb) == 6 * sizeof(float), "Aabb is lo xyz, hi xyz");
    HIPCHK(hipMemcpyAsync(ctx->d_inst_bounds, ctx->inst_bounds.data(), (size_t)N * sizeof(Aabb),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_itrav, ctx->itrav.data(), ctx->itrav.size() * sizeof(InstTrav),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_ishade, ctx->ishade.data(), ctx->ishade.size() * sizeof(InstShade),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(launch_tlas_build(ctx->stream, ctx->d_inst_bounds, N, ctx->d_tlas_ref, ctx->d_tlas, ctx->d_tblocks,
                             ctx->d_tlas_status));
    ctx->tlas_dirty = false;
    ctx->tlas_on_device = true;
    set_scene_box(ctx);
    return KPT_OK;
}

// Rebuilds the TLAS after instance edits, before the next use.
int ensure_tlas(kpt_ctx* ctx) {
    if (!ctx->tlas_dirty) return KPT_OK;
    return rebuild_tlas_device(ctx);
}

// The host copy of a device-built TLAS (export and statistics).
int fetch_tlas(kpt_ctx* ctx) {
    if (!ctx->tlas_on_device) return KPT_OK;
    const size_t n = ctx->tlas.nodes.size();
    int32_t built = 0;
    HIPCHK(hipMemcpyAsync(&built, ctx->d_tlas_status, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->tlas.nodes.data(), ctx->d_tlas_ref, n * sizeof(Node), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if ((size_t)built != n) return fail(ctx, KPT_ERR_HIP, "device TLAS build failed");
    ctx->tlas_on_device = false;
    return KPT_OK;
}
#define ENSURE_TLAS(c)                      \
    do {                                    \
        int rc_ = ensure_tlas(c);           \
        if (rc_ != KPT_OK) return rc_;      \
    } while (0)

FrameConst make_frame(const kpt_ctx* c, int32_t /*frame*/) {
    FrameConst fc;
    std::memset(&fc, 0, sizeof(fc));
    std::memcpy(fc.inv_view, c->cam.inv_view, 64);
    std::memcpy(fc.inv_proj, c->cam.inv_proj, 64);
    hm::Mat4 cvp = hm::mul(hm::ref_inverse(hm::load(c->cam.inv_proj)), hm::ref_inverse(hm::load(c->cam.inv_view)));
    hm::Mat4 pvp = hm::mul(hm::ref_inverse(hm::load(c->cam.prev_inv_proj)), hm::ref_inverse(hm::load(c->cam.prev_inv_view)));
    store_mat(cvp, fc.curr_vp);
    store_mat(pvp, fc.prev_vp);
    fc.cam_pos[0] = c->cam.inv_view[12];
    fc.cam_pos[1] = c->cam.inv_view[13];
    fc.cam_pos[2] = c->cam.inv_view[14];
    fc.prev_cam_pos[0] = c->cam.prev_inv_view[12];
    fc.prev_cam_pos[1] = c->cam.prev_inv_view[13];
    fc.prev_cam_pos[2] = c->cam.prev_inv_view[14];
    fc.temporal = c->settings.use_temporal_accumulation ? 1 : 0;
    for (int k = 0; k < 3; ++k) {
        const float lo = c->scene_lo[k];
        const float hi = c->scene_hi[k];
        fc.scene_lo[k] = lo;
        fc.scene_inv[k] = (hi > lo) ? 1.0f / (hi - lo) : 1.0f;
    }
    const kpt_settings& s = c->settings;
    // getSunDirection (integrator.cuh:222-230)
    float phi = s.sun_phi_rad, th = s.sun_theta_rad;
    v3 sun = normalize(mk3(dm::cosf(phi) * dm::cosf(th), dm::sinf(th), dm::sinf(phi) * dm::cosf(th)));
    fc.sun_dir[0] = sun.x; fc.sun_dir[1] = sun.y; fc.sun_dir[2] = sun.z;
    fc.sun_emission = s.sun_emission_factor;
    fc.sun_diam = s.sun_angular_diameter_rad;
    // The sky radiance toward the sun that sampleLdSun evaluates per call
    // (integrator.cuh:169) does not depend on the path: computed once per frame
    // by the same routine.
    v3 srad = sky_li(observer_pos(), sun, sun, s.sun_emission_factor);
    fc.sun_radiance[0] = srad.x; fc.sun_radiance[1] = srad.y; fc.sun_radiance[2] = srad.z;
    // angularDiameterToPhysicalDiameter(...) / 2 (atmosphere.cuh:79-82, integrator.cuh:153)
    fc.sun_radius = (2.0f * KPT_SOLAR_DIST * dm::tanf(s.sun_angular_diameter_rad / 2.0f)) / 2.0f;
    fc.min_sun_cos = dm::cosf(s.sun_angular_diameter_rad / 2.0f);
    fc.width = c->width;
    fc.height = c->height;
    fc.max_depth = s.max_depth;
    fc.seed_mode = s.seed_mode;
    fc.num_lights = (int32_t)c->lights.size();
    fc.num_instances = (int32_t)c->itrav.size();
    fc.tlas_root = c->tlas.root;
    return fc;
}

void free_frame(kpt_ctx* c) {
    dev_free(&c->fb.accum); dev_free(&c->fb.render); dev_free(&c->fb.gbuffer);
    dev_free(&c->fb.prev_gbuffer); dev_free(&c->fb.vbuffer); dev_free(&c->fb.output);
    for (int k = 0; k < MAX_MIPS; ++k) dev_free(&c->pb.mip[k]);
    dev_free(&c->pb.back);
    dev_free(&c->pb.up);
    c->pb.levels = 0;
}

// Renderer state before any setExposure call: AutoExposureProgram members
// (autoexposure_program.hpp:75-78, ExposureValues defaults pod_types.hpp:26-32)
// and the Film defaults (film.cuh:188-190).
// ---------------------------------------------------------------- radiance cache
void free_cache(kpt_ctx* c) {
    CacheBuffers& b = c->cb;
    dev_free(&b.w); dev_free(&b.adam_m); dev_free(&b.adam_v); dev_free(&b.grad); dev_free(&b.wb); dev_free(&b.wbt);
    dev_free(&b.partial); dev_free(&b.partial_sum); dev_free(&b.loss_partial); dev_free(&b.stats); dev_free(&b.q_pos); dev_free(&b.q_dir);
    dev_free(&b.q_beta); dev_free(&b.t_pos); dev_free(&b.t_dir); dev_free(&b.t_beta); dev_free(&b.t_lbefore);
    dev_free(&b.t_target); dev_free(&b.counters);
    b = CacheBuffers{};
}

void default_cache_settings(kpt_cache_settings* s) {
    std::memset(s, 0, sizeof(*s));
    s->train_samples = 65536;
    s->train_steps = 4;
    s->learning_rate = 1e-3f;
    s->radiance_scale = 1e4f;
    s->seed = 1;
    s->nee_at_vertex = 1;
    s->wide_roughness = 0.0f;
}

// The parameter buffers (allocated once) and the training workspace for
// batches of up to `batch` samples.
int cache_alloc_params(kpt_ctx* ctx, int32_t batch) {
    CacheBuffers& b = ctx->cb;
    if (!b.w) {
        HIPCHK(hipMalloc((void**)&b.w, RC_PARAMS * sizeof(float)));
        HIPCHK(hipMalloc((void**)&b.adam_m, RC_PARAMS * sizeof(float)));
        HIPCHK(hipMalloc((void**)&b.adam_v, RC_PARAMS * sizeof(float)));
        HIPCHK(hipMalloc((void**)&b.grad, RC_PARAMS * sizeof(float)));
        HIPCHK(hipMalloc((void**)&b.wb, RC_PADDED * sizeof(uint16_t)));
        HIPCHK(hipMalloc((void**)&b.wbt, RC_PADDED * sizeof(uint16_t)));
        HIPCHK(hipMalloc((void**)&b.stats, 4 * sizeof(float)));
        HIPCHK(hipMalloc((void**)&b.counters, 4 * sizeof(int32_t)));
        HIPCHK(hipMalloc((void**)&b.partial_sum, (size_t)RC_CHUNKS * RC_PADDED * sizeof(float)));
        HIPCHK(hipMemsetAsync(b.grad, 0, RC_PARAMS * sizeof(float), ctx->stream));
        HIPCHK(hipMemsetAsync(b.stats, 0, 4 * sizeof(float), ctx->stream));
        HIPCHK(hipMemsetAsync(b.counters, 0, 4 * sizeof(int32_t), ctx->stream));
    }
    const int need = cache_train_workgroups(std::max(batch, 1));
    if (need > b.max_wg) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        dev_free(&b.partial);
        dev_free(&b.loss_partial);
        HIPCHK(hipMalloc((void**)&b.partial, (size_t)need * RC_PADDED * sizeof(float)));
        HIPCHK(hipMalloc((void**)&b.loss_partial, (size_t)need * sizeof(float)));
        b.max_wg = need;
    }
    return KPT_OK;
}

// Per-frame records: one query per path, `train_cap` training records.
int cache_alloc_records(kpt_ctx* ctx, int64_t npx, int64_t train_cap) {
    CacheBuffers& b = ctx->cb;
    if (b.query_cap < npx) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        dev_free(&b.q_pos); dev_free(&b.q_dir); dev_free(&b.q_beta);
        HIPCHK(hipMalloc((void**)&b.q_pos, (size_t)npx * sizeof(float4)));
        HIPCHK(hipMalloc((void**)&b.q_dir, (size_t)npx * sizeof(float4)));
        HIPCHK(hipMalloc((void**)&b.q_beta, (size_t)npx * sizeof(float4)));
        b.query_cap = npx;
    }
    if (b.train_cap < train_cap) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        dev_free(&b.t_pos); dev_free(&b.t_dir); dev_free(&b.t_beta); dev_free(&b.t_lbefore); dev_free(&b.t_target);
        HIPCHK(hipMalloc((void**)&b.t_pos, (size_t)train_cap * sizeof(float4)));
        HIPCHK(hipMalloc((void**)&b.t_dir, (size_t)train_cap * sizeof(float4)));
        HIPCHK(hipMalloc((void**)&b.t_beta, (size_t)train_cap * sizeof(float4)));
        HIPCHK(hipMalloc((void**)&b.t_lbefore, (size_t)train_cap * sizeof(float4)));
        HIPCHK(hipMalloc((void**)&b.t_target, (size_t)train_cap * sizeof(float4)));
        b.train_cap = train_cap;
    }
    return KPT_OK;
}

// Uploads FP32 master weights, clears the Adam state, republishes BF16.
int cache_upload_weights(kpt_ctx* ctx, const float* w) {
    CacheBuffers& b = ctx->cb;
    HIPCHK(hipMemcpyAsync(b.w, w, RC_PARAMS * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(b.adam_m, 0, RC_PARAMS * sizeof(float), ctx->stream));
    HIPCHK(hipMemsetAsync(b.adam_v, 0, RC_PARAMS * sizeof(float), ctx->stream));
    HIPCHK(launch_cache_publish(ctx->stream, b));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->cache_steps = 0;
    ctx->cache_weights = true;
    return KPT_OK;
}

// He-uniform initialisation: w = (2u - 1) * sqrt(6 / fan_in), u from a PCG hash
// of (seed, parameter index).
std::vector<float> cache_init_weights(uint32_t seed) {
    std::vector<float> w(RC_PARAMS);
    for (int L = 0; L < RC_LAYERS; ++L) {
        const float a = std::sqrt(6.0f / (float)rc_in(L));
        const int n = rc_out(L) * rc_in(L);
        for (int k = 0; k < n; ++k) {
            const int i = rc_off(L) + k;
            const uint32_t h = pcg_hash(seed ^ pcg_hash((uint32_t)i + 0x9E3779B9u));
            const float u = (float)(h >> 8) * (1.0f / 16777216.0f);
            w[i] = (2.0f * u - 1.0f) * a;
        }
    }
    return w;
}

// Position box and output unit of the encoding.
CacheEncode cache_encode(const kpt_ctx* ctx) {
    CacheEncode e{};
    const kpt_cache_settings& s = ctx->cache_set;
    float lo[3], hi[3];
    bool box = true;
    for (int k = 0; k < 3; ++k) box = box && (s.bounds_max[k] > s.bounds_min[k]);
    for (int k = 0; k < 3; ++k) {
        if (box) { lo[k] = s.bounds_min[k]; hi[k] = s.bounds_max[k]; }
        else if (!ctx->tlas.nodes.empty()) { lo[k] = ctx->tlas.nodes[0].lo[k]; hi[k] = ctx->tlas.nodes[0].hi[k]; }
        else { lo[k] = 0.0f; hi[k] = 1.0f; }
        const float ext = hi[k] - lo[k];
        e.bmin[k] = lo[k];
        e.binv[k] = ext > 0.0f ? 1.0f / ext : 1.0f;
    }
    e.scale = s.radiance_scale;
    return e;
}

void default_exposure(ExposureState* e) {
    std::memset(e, 0, sizeof(*e));
    e->ae.max_iso = 6400;
    e->ae.min_iso = 100;
    e->ae.min_shutter_secs = 0.00025f;
    e->ae.max_shutter_secs = 0.0333f;
    e->scalar = 1.0f;
    e->white_ev = 6.5f;
    e->black_ev = -10.0f;
    e->avg_lum = 0.0f;
}

void free_wf_state(WavefrontState& w) {
    dev_free(&w.ray_o); dev_free(&w.ray_d); dev_free(&w.beta); dev_free(&w.L); dev_free(&w.prev_pos);
    dev_free(&w.vis_a); dev_free(&w.vis_n); dev_free(&w.vis_p); dev_free(&w.hit); dev_free(&w.hit_v);
    dev_free(&w.queue[0]); dev_free(&w.queue[1]); dev_free(&w.shadow_queue); dev_free(&w.shadow_o);
    dev_free(&w.shadow_d); dev_free(&w.shadow_c); dev_free(&w.shadow_z);
    dev_free(&w.sort_tmp); dev_free(&w.counters); dev_free(&w.hist); dev_free(&w.tpool);
    w.tpool_waves = 0;
    for (int k = 0; k < 2; ++k) { dev_free(&w.rkeys[k]); dev_free(&w.rvals[k]); }
    if (w.rtemp) { (void)hipFree(w.rtemp); w.rtemp = nullptr; }
    w.rtemp_bytes = 0;
    w.capacity = 0;
}

void free_wavefront(kpt_ctx* c) {
    free_wf_state(c->ws);
    free_wf_state(c->ws2);
}

// Path state for n paths in w (ctx->ws, or ctx->ws2 for a second frame in flight).
int alloc_wf_state(kpt_ctx* ctx, WavefrontState& w, int64_t n) {
    if (w.capacity >= n) return KPT_OK;
    free_wf_state(w);
    size_t f4 = (size_t)n * sizeof(float4);
    size_t i4 = (size_t)n * sizeof(int32_t);
    HIPCHK(hipMalloc((void**)&w.ray_o, f4)); HIPCHK(hipMalloc((void**)&w.ray_d, f4));
    HIPCHK(hipMalloc((void**)&w.beta, f4)); HIPCHK(hipMalloc((void**)&w.L, f4));
    HIPCHK(hipMalloc((void**)&w.prev_pos, f4)); HIPCHK(hipMalloc((void**)&w.vis_a, f4));
    HIPCHK(hipMalloc((void**)&w.vis_n, f4)); HIPCHK(hipMalloc((void**)&w.vis_p, f4));
    HIPCHK(hipMalloc((void**)&w.hit, f4)); HIPCHK(hipMalloc((void**)&w.hit_v, (size_t)n * sizeof(float)));
    HIPCHK(hipMalloc((void**)&w.queue[0], i4)); HIPCHK(hipMalloc((void**)&w.queue[1], i4));
    HIPCHK(hipMalloc((void**)&w.shadow_queue, i4));
    HIPCHK(hipMalloc((void**)&w.shadow_o, f4)); HIPCHK(hipMalloc((void**)&w.shadow_d, f4));
    HIPCHK(hipMalloc((void**)&w.shadow_c, f4)); HIPCHK(hipMalloc((void**)&w.shadow_z, f4));
    HIPCHK(hipMalloc((void**)&w.sort_tmp, i4));
    HIPCHK(hipMalloc((void**)&w.counters, (size_t)WF_COUNTER_WORDS * sizeof(int32_t)));
    // parked tail traces: the waves of the default tail grid (8 workgroups of
    // 256 lanes per CU), 256 bytes per slot; larger grids park in these only
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                                                       dev) != hipSuccess || cus <= 0)
            cus = 256;
        const int64_t waves = (int64_t)cus * 8 * 4;
        HIPCHK(hipMalloc((void**)&w.tpool, (size_t)waves * TAIL_POOL_BYTES));
        w.tpool_waves = waves;
    }
    HIPCHK(hipMalloc((void**)&w.hist, (size_t)64 * ((n + 255) / 256) * sizeof(int32_t)));
    for (int k = 0; k < 2; ++k) {
        HIPCHK(hipMalloc((void**)&w.rkeys[k], i4));
        HIPCHK(hipMalloc((void**)&w.rvals[k], i4));
    }
    w.rtemp_bytes = ray_sort_temp_bytes(n);
    HIPCHK(hipMalloc(&w.rtemp, std::max<size_t>(w.rtemp_bytes, 256)));
    w.capacity = n;
    return KPT_OK;
}

int alloc_wavefront(kpt_ctx* ctx, int64_t n) { return alloc_wf_state(ctx, ctx->ws, n); }

}  // namespace

// Captures what enqueue() launches on the context stream (relaxed capture
// mode) and replays it as one graph: the first capture is instantiated into
// *exec, later ones update it in place (hipGraphExecUpdate; re-instantiated
// when the topology changed).  Nothing enqueued here may be a memset node:
// counters and queues are cleared by fill kernels.
template <class Fn>
static hipError_t replay_captured(kpt_ctx* ctx, hipGraphExec_t* exec, Fn&& enqueue) {
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return e;
    e = enqueue();
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess && *exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(*exec, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(*exec);
            *exec = nullptr;
        }
    }
    if (e == hipSuccess && !*exec) e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipGraphLaunch(*exec, ctx->stream);
    if (g) (void)hipGraphDestroy(g);
    return e;
}

// Scattered 16-byte reads with a known byte count (kpt_measure_gather16).
__global__ __launch_bounds__(256) void k_gather16(const float4* src, int64_t src_records, const int32_t* idx,
                                                  int64_t n, float4* dst) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int32_t k = idx[i];
    dst[i] = src[(k >= 0 && k < src_records) ? k : 0];
}

extern "C" {

int kpt_get_abi_version(void) { return KPT_ABI_VERSION; }

int kpt_ctx_create(int device, kpt_ctx** out) {
    if (!out) return KPT_ERR_INVALID_ARG;
    *out = nullptr;
    kpt_ctx* ctx = new kpt_ctx();
    default_settings(&ctx->settings);
    default_exposure(&ctx->h_expo);
    default_cache_settings(&ctx->cache_set);
    hm::Mat4 I = hm::identity();
    std::memcpy(ctx->cam.inv_view, I.m, 64); std::memcpy(ctx->cam.inv_proj, I.m, 64);
    std::memcpy(ctx->cam.prev_inv_view, I.m, 64); std::memcpy(ctx->cam.prev_inv_proj, I.m, 64);
    if (device == -2) {  // host-only context: builders and queries, no GPU work
        ctx->host_only = true;
        *out = ctx;
        return KPT_OK;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        delete ctx;
        return KPT_ERR_NO_DEVICE;
    }
    if (device < 0) { if (hipGetDevice(&device) != hipSuccess) device = 0; }
    if (device >= count) { delete ctx; return KPT_ERR_INVALID_ARG; }
    ctx->device = device;
    ctx->blas_device = 1;                // same trees as the host builder, built faster on the device
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&ctx->d_overflow, sizeof(int32_t)) != hipSuccess ||
        hipMemset(ctx->d_overflow, 0, sizeof(int32_t)) != hipSuccess || hipEventCreate(&ctx->ev0) != hipSuccess ||
        hipEventCreate(&ctx->ev1) != hipSuccess || hipEventCreate(&ctx->tev0) != hipSuccess ||
        hipEventCreate(&ctx->tev1) != hipSuccess ||
        hipMalloc((void**)&ctx->d_counts, 16 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void**)&ctx->pb.expo, sizeof(ExposureState)) != hipSuccess ||
        hipMalloc((void**)&ctx->pb.hist, 256 * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(ctx->pb.hist, 0, 256 * sizeof(unsigned long long)) != hipSuccess) {
        delete ctx;
        return KPT_ERR_HIP;
    }
    if (scratch_address(device, &ctx->scratch_addr) != hipSuccess) {
        delete ctx;
        return KPT_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cus = prop.multiProcessorCount;
    if (hipMemcpy(ctx->pb.expo, &ctx->h_expo, sizeof(ExposureState), hipMemcpyHostToDevice) != hipSuccess) {
        delete ctx;
        return KPT_ERR_HIP;
    }
    for (int r = 0; r < kpt_ctx::INS_RING; ++r) {
        for (int k = 0; k < WF_MAX_TRACE_LAUNCHES; ++k) {
            if (hipEventCreate(&ctx->ins[r].begin[k]) != hipSuccess ||
                hipEventCreate(&ctx->ins[r].end[k]) != hipSuccess) {
                delete ctx;
                return KPT_ERR_HIP;
            }
        }
        bool xok = true;
        for (int x = 0; x < WF_XSTAGES; ++x)
            xok = xok && hipEventCreate(&ctx->ins[r].xbegin[x]) == hipSuccess &&
                  hipEventCreate(&ctx->ins[r].xend[x]) == hipSuccess;
        if (!xok || hipEventCreate(&ctx->ins[r].tail_begin) != hipSuccess ||
            hipEventCreate(&ctx->ins[r].tail_end) != hipSuccess) {
            delete ctx;
            return KPT_ERR_HIP;
        }
    }
    *out = ctx;
    return KPT_OK;
}

int kpt_ctx_destroy(kpt_ctx* ctx) {
    if (!ctx) return KPT_ERR_INVALID_ARG;
    if (!ctx->host_only) {
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
        for (int k = 0; k < 2; ++k)
            if (ctx->lane_done[k]) (void)hipEventDestroy(ctx->lane_done[k]);
        if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
        dev_free(&ctx->d_blas); dev_free(&ctx->d_tlas); dev_free(&ctx->d_tris); dev_free(&ctx->d_attrs);
        dev_free(&ctx->d_bblocks); dev_free(&ctx->d_tblocks);
        dev_free(&ctx->d_itrav); dev_free(&ctx->d_ishade); dev_free(&ctx->d_mat); dev_free(&ctx->d_tex);
        dev_free(&ctx->d_pix); dev_free(&ctx->d_lights); dev_free(&ctx->d_lverts); dev_free(&ctx->d_overflow);
        free_frame(ctx);
        free_wavefront(ctx);
        if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
        if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
        if (ctx->tev0) (void)hipEventDestroy(ctx->tev0);
        if (ctx->tev1) (void)hipEventDestroy(ctx->tev1);
        for (int r = 0; r < kpt_ctx::INS_RING; ++r) {
            for (int k = 0; k < WF_MAX_TRACE_LAUNCHES; ++k) {
                if (ctx->ins[r].begin[k]) (void)hipEventDestroy(ctx->ins[r].begin[k]);
                if (ctx->ins[r].end[k]) (void)hipEventDestroy(ctx->ins[r].end[k]);
            }
            if (ctx->ins[r].tail_begin) (void)hipEventDestroy(ctx->ins[r].tail_begin);
            if (ctx->ins[r].tail_end) (void)hipEventDestroy(ctx->ins[r].tail_end);
            for (int x = 0; x < WF_XSTAGES; ++x) {
                if (ctx->ins[r].xbegin[x]) (void)hipEventDestroy(ctx->ins[r].xbegin[x]);
                if (ctx->ins[r].xend[x]) (void)hipEventDestroy(ctx->ins[r].xend[x]);
            }
        }
        if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
        if (ctx->views_graph_exec) (void)hipGraphExecDestroy(ctx->views_graph_exec);
        dev_free(&ctx->d_counts);
        dev_free(&ctx->d_heat);
        dev_free(&ctx->d_inst_bounds);
        dev_free(&ctx->d_tlas_ref);
        dev_free(&ctx->d_tlas_status);
        dev_free(&ctx->d_tlas_parent);
        dev_free(&ctx->d_tlas_arrivals);
        dev_free(&ctx->pb.expo);
        dev_free(&ctx->pb.hist);
        dev_free(&ctx->d_views);
        free_cache(ctx);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return KPT_OK;
}

const char* kpt_last_error(const kpt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int kpt_upload_scene(kpt_ctx* ctx, const kpt_scene_desc* d) {
    if (!ctx || !d) return KPT_ERR_INVALID_ARG;
    const int32_t T = d->num_triangles;
    if (T < 0 || d->num_meshes < 0 || d->num_materials < 0 || d->num_textures < 0)
        return fail(ctx, KPT_ERR_INVALID_ARG, "negative count");
    if (T > 0 && (!d->positions || !d->normals || !d->uvs || !d->material_ids))
        return fail(ctx, KPT_ERR_INVALID_ARG, "missing triangle arrays");
    // meshes must tile the triangle array in order (the reference's builder
    // indexes its centroid cache by global triangle id, as_builder.cuh:238-248)
    int32_t expect = 0;
    for (int32_t m = 0; m < d->num_meshes; ++m) {
        if (d->mesh_first[m] != expect || d->mesh_count[m] < 0)
            return fail(ctx, KPT_ERR_INVALID_ARG, "meshes must be contiguous and in triangle order");
        expect += d->mesh_count[m];
    }
    if (expect != T) return fail(ctx, KPT_ERR_INVALID_ARG, "mesh triangle counts do not sum to num_triangles");
    for (int32_t t = 0; t < T; ++t)
        if (d->material_ids[t] < 0 || d->material_ids[t] >= d->num_materials)
            return fail(ctx, KPT_ERR_INVALID_ARG, "triangle material id out of range");
    const int32_t NI = d->num_instances > 0 ? d->num_instances : d->num_meshes;
    if (d->num_instances == 0 && d->num_meshes > 0 && !d->mesh_model)
        return fail(ctx, KPT_ERR_INVALID_ARG, "mesh_model required when num_instances == 0");
    for (int32_t i = 0; i < d->num_instances; ++i)
        if (d->instance_mesh[i] < 0 || d->instance_mesh[i] >= d->num_meshes)
            return fail(ctx, KPT_ERR_INVALID_ARG, "instance mesh index out of range");
    for (int32_t i = 0; i < d->num_textures; ++i) {
        const kpt_texture_desc& t = d->textures[i];
        uint64_t need = (uint64_t)t.width * t.height * t.channels;
        if (t.offset + need > d->num_pixel_bytes || t.width == 0 || t.height == 0 || t.channels == 0)
            return fail(ctx, KPT_ERR_INVALID_ARG, "texture outside the pixel buffer");
    }
    // texture ids of materials must name a texture or be -1 (the reference
    // indexes its texture buffer unchecked, material.cu:23-48)
    for (int32_t i = 0; i < d->num_materials; ++i) {
        const kpt_material& m = d->materials[i];
        const int32_t ids[5] = {m.albedo_texture_id, m.orm_texture_id, m.transmission_texture_id,
                                m.emission_texture_id, m.normal_texture_id};
        for (int k = 0; k < 5; ++k)
            if (ids[k] < -1 || ids[k] >= d->num_textures)
                return fail(ctx, KPT_ERR_INVALID_ARG, "material texture id out of range");
    }
    ctx->scene_ready = false;
    ctx->pos.assign(d->positions, d->positions + 9 * (size_t)T);
    ctx->uv.assign(d->uvs, d->uvs + 6 * (size_t)T);
    ctx->mat_ids.assign(d->material_ids, d->material_ids + T);
    ctx->mesh_first.assign(d->mesh_first, d->mesh_first + d->num_meshes);
    ctx->mesh_count.assign(d->mesh_count, d->mesh_count + d->num_meshes);
    ctx->inst_mesh.resize(NI);
    ctx->inst_model.resize(16 * (size_t)NI);
    for (int32_t i = 0; i < NI; ++i) {
        ctx->inst_mesh[i] = d->num_instances > 0 ? d->instance_mesh[i] : i;
        const float* src = d->num_instances > 0 ? d->instance_model + 16 * (size_t)i : d->mesh_model + 16 * (size_t)i;
        std::memcpy(&ctx->inst_model[16 * (size_t)i], src, 64);
    }
    // materials (Material ctor from MaterialSceneEntity, renderer.cu:372-382)
    ctx->materials.resize(d->num_materials);
    for (int32_t i = 0; i < d->num_materials; ++i) {
        const kpt_material& s = d->materials[i];
        kpt::Material& m = ctx->materials[i];
        std::memset(&m, 0, sizeof(m));
        m.albedo_tex = s.albedo_texture_id;
        for (int k = 0; k < 3; ++k) { m.albedo[k] = s.albedo[k]; m.emission[k] = s.emission_factor[k]; }
        m.orm_tex = s.orm_texture_id; m.metallic = s.metallic; m.roughness = s.roughness;
        m.transmission_tex = s.transmission_texture_id; m.transmission = s.transmission;
        m.ior = s.ior; m.emission_tex = s.emission_texture_id; m.emission_scale = s.emission_scale_nits;
        m.normal_tex = s.normal_texture_id; m.normal_scale = s.normal_scale;
    }
    ctx->textures.resize(d->num_textures);
    for (int32_t i = 0; i < d->num_textures; ++i) {
        kpt::Texture& t = ctx->textures[i];
        std::memset(&t, 0, sizeof(t));
        t.width = (int32_t)d->textures[i].width; t.height = (int32_t)d->textures[i].height;
        t.channels = (int32_t)d->textures[i].channels; t.bit_depth = 8; t.offset = (int64_t)d->textures[i].offset;
    }
    ctx->num_pixel_bytes = (int64_t)d->num_pixel_bytes;
    // BLAS per mesh: host threads, or all meshes at once on the device
    ctx->blas.assign(d->num_meshes, BlasTree());
    const auto build_t0 = std::chrono::steady_clock::now();
    // The device builder needs a device, at most BLAS_DEVICE_MAX_TRIS
    // triangles and finite positions (its bin keys order finite floats); a
    // scene outside that, or a device build that fails, is built by the host
    // threads instead (the same trees; the reference builds any scene on the
    // host, as_builder.cuh:26-35).
    bool on_device = ctx->blas_device && T > 0 && !ctx->host_only && T <= BLAS_DEVICE_MAX_TRIS;
    for (size_t i = 0; on_device && i < ctx->pos.size(); ++i)
        if (!std::isfinite(ctx->pos[i])) on_device = false;
    if (on_device) {
        std::vector<Node> dn;
        std::vector<int32_t> dorder, dused, ddepth;
        hipError_t be = device_build_blas(ctx->stream, ctx->pos.data(), ctx->mesh_first.data(), ctx->mesh_count.data(),
                                          d->num_meshes, T, &dn, &dorder, &dused, &ddepth);
        if (be != hipSuccess) {
            (void)hipGetLastError();
            on_device = false;
        } else {
            size_t base = 0;
            for (int32_t m = 0; m < d->num_meshes; ++m) {
                const int32_t cnt = ctx->mesh_count[m];
                if (cnt <= 0) continue;
                BlasTree& bt = ctx->blas[m];
                const size_t cap = 2 * (size_t)cnt - 1;
                bt.nodes.assign(dn.begin() + base, dn.begin() + base + cap);
                bt.order.assign(dorder.begin() + ctx->mesh_first[m], dorder.begin() + ctx->mesh_first[m] + cnt);
                bt.used = dused[m];
                bt.max_depth = ddepth[m];
                for (int k = 0; k < 3; ++k) { bt.bounds.lo[k] = bt.nodes[0].lo[k]; bt.bounds.hi[k] = bt.nodes[0].hi[k]; }
                base += cap;
            }
        }
    }
    if (!on_device) {
        std::vector<std::thread> pool;
        const float* P = ctx->pos.data();
        for (int32_t m = 0; m < d->num_meshes; ++m)
            pool.emplace_back([ctx, P, m]() { ctx->blas[m] = build_blas(P, ctx->mesh_first[m], ctx->mesh_count[m]); });
        for (auto& th : pool) th.join();
    }
    ctx->blas_build_ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - build_t0).count();
    ctx->blas_build_ms[1] = on_device ? 1.0 : 0.0;
    // global node array: each root at an odd index so every sibling pair
    // starts at an even index (one 64-byte line)
    ctx->nodes.clear();
    ctx->geom_base.assign(d->num_meshes, -1);
    for (int32_t m = 0; m < d->num_meshes; ++m) {
        const BlasTree& b = ctx->blas[m];
        if (b.used == 0) continue;
        if ((ctx->nodes.size() & 1) == 0) { Node pad; std::memset(&pad, 0, sizeof(pad)); ctx->nodes.push_back(pad); }
        int32_t base = (int32_t)ctx->nodes.size();
        ctx->geom_base[m] = base;
        for (int32_t k = 0; k < b.used; ++k) {
            Node n = b.nodes[k];
            if (n.b == 0u) n.a = (uint32_t)(base + (int32_t)n.a);
            else n.a = (uint32_t)(ctx->mesh_first[m] + (int32_t)n.a);
            ctx->nodes.push_back(n);
        }
        if (ctx->nodes.size() & 1) { Node pad; std::memset(&pad, 0, sizeof(pad)); ctx->nodes.push_back(pad); }
    }
    // triangles in slot order, attributes in primitive order
    std::vector<TriPos> tris(T);
    std::vector<TriAttr> attrs(T);
    for (int32_t m = 0; m < d->num_meshes; ++m) {
        const BlasTree& b = ctx->blas[m];
        for (int32_t s = 0; s < ctx->mesh_count[m]; ++s) {
            int32_t prim = b.order[s];
            const float* p = &ctx->pos[9 * (size_t)prim];
            TriPos& tp = tris[ctx->mesh_first[m] + s];
            for (int k = 0; k < 3; ++k) { tp.v0[k] = p[k]; tp.e1[k] = p[3 + k] - p[k]; tp.e2[k] = p[6 + k] - p[k]; }
            tp.prim = prim; tp.pad0 = 0.0f; tp.pad1 = 0.0f;
        }
    }
    // lights: emissive materials (renderer.cu:391-401), placed by the first
    // instance of their mesh
    std::vector<int32_t> first_inst(d->num_meshes, -1);
    for (int32_t i = 0; i < NI; ++i) if (first_inst[ctx->inst_mesh[i]] < 0) first_inst[ctx->inst_mesh[i]] = i;
    ctx->lights.clear();
    ctx->light_verts.clear();
    for (int32_t t = 0; t < T; ++t) {
        const float* p = &ctx->pos[9 * (size_t)t];
        const float* n = d->normals + 9 * (size_t)t;
        TriAttr& a = attrs[t];
        // Triangle constructor (triangle.cuh:54-64): geometric normal flipped
        // toward the mean vertex normal, then normalised
        v3 p0 = mk3(p[0], p[1], p[2]), p1 = mk3(p[3], p[4], p[5]), p2 = mk3(p[6], p[7], p[8]);
        v3 gn = cross(p1 - p0, p2 - p0);
        v3 avg = (mk3(n[0], n[1], n[2]) + mk3(n[3], n[4], n[5]) + mk3(n[6], n[7], n[8])) / 3.0f;
        float dd = dot(gn, avg);
        v3 lg = normalize((dd < 0.0f) ? -gn : gn);
        a.lgn[0] = lg.x; a.lgn[1] = lg.y; a.lgn[2] = lg.z;
        a.material = ctx->mat_ids[t];
        const float* u = &ctx->uv[6 * (size_t)t];
        a.uv0[0] = u[0]; a.uv0[1] = u[1]; a.uv1[0] = u[2]; a.uv1[1] = u[3]; a.uv2[0] = u[4]; a.uv2[1] = u[5];
        a.light = -1;
        a.inst_of_light = -1;
        const kpt::Material& mat = ctx->materials[a.material];
        v3 e = mk3(mat.emission[0], mat.emission[1], mat.emission[2]) * mat.emission_scale;
        if (e.x > 0.0f || e.y > 0.0f || e.z > 0.0f) {
            int32_t mesh = (int32_t)(std::upper_bound(ctx->mesh_first.begin(), ctx->mesh_first.end(), t) - ctx->mesh_first.begin()) - 1;
            kpt::Light l;
            l.area = 0.5f * length(cross(p1 - p0, p2 - p0));
            l.prim = t;
            l.inst = first_inst[mesh];
            l.pad = 0;
            if (l.inst < 0) continue;   // mesh never instanced: not a light
            a.light = (int32_t)ctx->lights.size();
            a.inst_of_light = l.inst;
            ctx->lights.push_back(l);
            ctx->light_verts.insert(ctx->light_verts.end(), p, p + 9);
        }
    }
    // instances + TLAS
    ctx->itrav.assign(NI, InstTrav());
    ctx->ishade.assign(NI, InstShade());
    ctx->inst_bounds.assign(NI, Aabb());
    for (int32_t i = 0; i < NI; ++i) make_instance(ctx, i, &ctx->inst_model[16 * (size_t)i], false);
    if (!ctx->host_only) {
        HIPCHK(hipSetDevice(ctx->device));
        HIPCHK(dev_upload(&ctx->d_blas, ctx->nodes));
        std::vector<NodeBlock> bb;
        if (!make_blocks(ctx->nodes, [](const Node& n) { return n.b == 0u; }, &bb))
            return fail(ctx, KPT_ERR_UNSUPPORTED, "BLAS sibling pair not at an even node index");
        if (ctx->nodes.size() >= (size_t)ENTRY_PAYLOAD)
            return fail(ctx, KPT_ERR_UNSUPPORTED, "BLAS node count exceeds the traversal entry range");
        HIPCHK(dev_upload(&ctx->d_bblocks, bb));
        HIPCHK(dev_upload(&ctx->d_tris, tris));
        HIPCHK(dev_upload(&ctx->d_attrs, attrs));
        HIPCHK(dev_upload(&ctx->d_mat, ctx->materials));
        HIPCHK(dev_upload(&ctx->d_tex, ctx->textures));
        // Texture::evaluate reads three bytes per texel whatever the channel
        // count (texture.cuh:54-58): pad so a 1- or 2-channel last texel stays
        // inside the allocation; the bounds checks still use num_pixel_bytes.
        std::vector<uint8_t> pix(d->num_pixel_bytes + 16, 0);
        if (d->num_pixel_bytes) std::memcpy(pix.data(), d->pixels, d->num_pixel_bytes);
        HIPCHK(dev_upload(&ctx->d_pix, pix));
        HIPCHK(dev_upload(&ctx->d_lights, ctx->lights));
        HIPCHK(dev_upload(&ctx->d_lverts, ctx->light_verts));
    }
    int rc = upload_tlas(ctx);
    if (rc != KPT_OK) return rc;
    ctx->scene_ready = true;
    ctx->frame_index = 0;
    return KPT_OK;
}

int kpt_set_instance_transforms(kpt_ctx* ctx, int32_t first, int32_t count, const float* models) {
    if (!ctx || !models || count < 0) return KPT_ERR_INVALID_ARG;
    if (!ctx->scene_ready) return fail(ctx, KPT_ERR_STATE, "no scene");
    if (first < 0 || (int64_t)first + count > (int64_t)ctx->itrav.size())
        return fail(ctx, KPT_ERR_INVALID_ARG, "instance out of range");
    for (int32_t k = 0; k < count; ++k) {
        const float* m = models + 16 * (size_t)k;
        std::memcpy(&ctx->inst_model[16 * (size_t)(first + k)], m, 64);
        make_instance(ctx, first + k, m, true);
    }
    // setMeshTransform resets accumulation unless temporal (renderer.cu:264-266)
    if (!ctx->settings.use_temporal_accumulation) {
        ctx->frame_index = 0;
        if (!ctx->host_only && ctx->fb.accum)
            HIPCHK(launch_clear(ctx->stream, ctx->fb.accum, (int64_t)ctx->width * ctx->height));
    }
    // the TLAS is rebuilt before the next frame or trace (updateTLAS, renderer.cu:115)
    ctx->tlas_dirty = true;
    return KPT_OK;
}

int kpt_set_instance_transform(kpt_ctx* ctx, int32_t instance, const float model[16]) {
    return kpt_set_instance_transforms(ctx, instance, 1, model);
}

int kpt_set_material(kpt_ctx* ctx, int32_t index, const kpt_material* s) {
    if (!ctx || !s) return KPT_ERR_INVALID_ARG;
    if (index < 0 || index >= (int32_t)ctx->materials.size()) return fail(ctx, KPT_ERR_INVALID_ARG, "material out of range");
    kpt::Material& m = ctx->materials[index];
    m.albedo_tex = s->albedo_texture_id;
    for (int k = 0; k < 3; ++k) { m.albedo[k] = s->albedo[k]; m.emission[k] = s->emission_factor[k]; }
    m.orm_tex = s->orm_texture_id; m.metallic = s->metallic; m.roughness = s->roughness;
    m.transmission_tex = s->transmission_texture_id; m.transmission = s->transmission;
    m.ior = s->ior; m.emission_tex = s->emission_texture_id; m.emission_scale = s->emission_scale_nits;
    m.normal_tex = s->normal_texture_id; m.normal_scale = s->normal_scale;
    ctx->frame_index = 0;   // renderer.cu:223
    if (ctx->host_only) return KPT_OK;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(dev_upload(&ctx->d_mat, ctx->materials));
    if (ctx->fb.accum) HIPCHK(launch_clear(ctx->stream, ctx->fb.accum, (int64_t)ctx->width * ctx->height));
    return KPT_OK;
}

int kpt_scene_info(const kpt_ctx* ctx, int64_t out[8]) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    int64_t md = 0;
    for (const auto& b : ctx->blas) md = std::max<int64_t>(md, b.max_depth);
    out[0] = (int64_t)ctx->pos.size() / 9;
    out[1] = (int64_t)ctx->nodes.size();
    out[2] = (int64_t)ctx->tlas.nodes.size();
    out[3] = (int64_t)ctx->itrav.size();
    out[4] = md;
    out[5] = ctx->tlas.max_depth;
    out[6] = (int64_t)ctx->lights.size();
    out[7] = (int64_t)ctx->materials.size();
    return KPT_OK;
}

int kpt_get_blas(const kpt_ctx* ctx, int32_t mesh, void* nodes32, int64_t cap_nodes, int32_t* tri_index,
                 int64_t cap_tris, int64_t out_counts[2]) {
    if (!ctx || !nodes32 || !tri_index || !out_counts) return KPT_ERR_INVALID_ARG;
    if (mesh < 0 || mesh >= (int32_t)ctx->blas.size()) return KPT_ERR_INVALID_ARG;
    const BlasTree& b = ctx->blas[mesh];
    if (b.used > cap_nodes || (int64_t)b.order.size() > cap_tris) return KPT_ERR_INVALID_ARG;
    struct RefNode { int32_t first; uint32_t count; float pmin[3]; float pmax[3]; };
    RefNode* out = (RefNode*)nodes32;
    for (int32_t k = 0; k < b.used; ++k) {
        const Node& n = b.nodes[k];
        out[k].first = (int32_t)n.a;
        out[k].count = n.b;
        for (int j = 0; j < 3; ++j) { out[k].pmin[j] = n.lo[j]; out[k].pmax[j] = n.hi[j]; }
    }
    for (size_t s = 0; s < b.order.size(); ++s) tri_index[s] = b.order[s];
    out_counts[0] = b.used;
    out_counts[1] = (int64_t)b.order.size();
    return KPT_OK;
}

int kpt_get_tlas(const kpt_ctx* ctx, void* nodes32, int64_t cap_nodes, int64_t* out_count) {
    if (!ctx || !nodes32 || !out_count) return KPT_ERR_INVALID_ARG;
    if (ctx->tlas_dirty) {
        int rc = ensure_tlas(const_cast<kpt_ctx*>(ctx));
        if (rc != KPT_OK) return rc;
    }
    {
        int rc = fetch_tlas(const_cast<kpt_ctx*>(ctx));
        if (rc != KPT_OK) return rc;
    }
    if ((int64_t)ctx->tlas.nodes.size() > cap_nodes) return KPT_ERR_INVALID_ARG;
    struct RefT { float pmin[3]; float pmax[3]; uint32_t lr; int32_t blas; };
    RefT* out = (RefT*)nodes32;
    for (size_t k = 0; k < ctx->tlas.nodes.size(); ++k) {
        const Node& n = ctx->tlas.nodes[k];
        for (int j = 0; j < 3; ++j) { out[k].pmin[j] = n.lo[j]; out[k].pmax[j] = n.hi[j]; }
        out[k].lr = n.a;
        out[k].blas = (int32_t)n.b;
    }
    *out_count = (int64_t)ctx->tlas.nodes.size();
    return KPT_OK;
}

// The traversal kernels raise the context's overflow flag when a lane's
// stack would pass LDS_STACK + SPILL_STACK entries (the lane's query then
// ends early).  Read after the stream has drained: a raised flag is cleared
// and reported once, by the first synchronising call after the launch
// (kpt_render, kpt_render_views, kpt_sync).
static int report_overflow(kpt_ctx* ctx) {
    int32_t ovf = 0;
    HIPCHK(hipMemcpy(&ovf, ctx->d_overflow, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (!ovf) return KPT_OK;
    HIPCHK(hipMemset(ctx->d_overflow, 0, sizeof(int32_t)));
    return fail(ctx, KPT_ERR_OVERFLOW, "traversal stack overflow (depth > 128)");
}

int kpt_trace_closest(kpt_ctx* ctx, const kpt_ray* rays, kpt_hit* hits, int64_t n) {
    if (!ctx || (n > 0 && (!rays || !hits)) || n < 0) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    if (!ctx->scene_ready) return fail(ctx, KPT_ERR_STATE, "no scene");
    ENSURE_TLAS(ctx);
    HIPCHK(launch_trace_closest(ctx->stream, scene_view(ctx), rays, hits, n, ctx->d_overflow, nullptr));
    return KPT_OK;
}

int kpt_trace_any(kpt_ctx* ctx, const kpt_ray* rays, int32_t* occluded, int64_t n) {
    if (!ctx || (n > 0 && (!rays || !occluded)) || n < 0) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    if (!ctx->scene_ready) return fail(ctx, KPT_ERR_STATE, "no scene");
    ENSURE_TLAS(ctx);
    HIPCHK(launch_trace_any(ctx->stream, scene_view(ctx), rays, occluded, n, ctx->d_overflow));
    return KPT_OK;
}

int kpt_trace_closest_counted(kpt_ctx* ctx, const kpt_ray* rays, kpt_hit* hits, int64_t n,
                              kpt_trace_counters* host_counters) {
    if (!ctx || !host_counters || n < 0 || (n > 0 && (!rays || !hits))) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    if (!ctx->scene_ready) return fail(ctx, KPT_ERR_STATE, "no scene");
    ENSURE_TLAS(ctx);
    kpt_trace_counters* d = nullptr;
    HIPCHK(hipMalloc((void**)&d, sizeof(kpt_trace_counters)));
    hipError_t e = hipMemsetAsync(d, 0, sizeof(kpt_trace_counters), ctx->stream);
    if (e == hipSuccess) e = launch_trace_closest(ctx->stream, scene_view(ctx), rays, hits, n, ctx->d_overflow, d);
    if (e == hipSuccess) e = hipMemcpyAsync(host_counters, d, sizeof(kpt_trace_counters), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(ctx, KPT_ERR_HIP, std::string("counted trace: ") + hipGetErrorString(e));
    host_counters->rays = (uint64_t)n;
    return KPT_OK;
}

int kpt_trace_closest_counted_noalloc(kpt_ctx* ctx, const kpt_ray* rays, kpt_hit* hits, int64_t n,
                                      kpt_trace_counters* host_counters) {
    if (!ctx || !host_counters || n < 0 || (n > 0 && (!rays || !hits))) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    if (!ctx->scene_ready) return fail(ctx, KPT_ERR_STATE, "no scene");
    // the TLAS of pending instance edits is rebuilt into freshly allocated
    // device buffers: that is left to an allocating call
    if (ctx->tlas_dirty)
        return fail(ctx, KPT_ERR_STATE,
                    "instance transforms changed: the TLAS rebuild allocates; call kpt_trace_closest or kpt_render first");
    if (!ctx->scratch_addr) return fail(ctx, KPT_ERR_STATE, "no device scratch on this context");
    // the counters start at zero: a fill kernel, not a memset node (callers may capture)
    HIPCHK(launch_clear(ctx->stream, reinterpret_cast<float4*>(ctx->scratch_addr),
                        (int64_t)(NOALLOC_SCRATCH_WORDS * sizeof(unsigned long long) / sizeof(float4))));
    HIPCHK(launch_trace_closest(ctx->stream, scene_view(ctx), rays, hits, n, ctx->d_overflow,
                                reinterpret_cast<kpt_trace_counters*>(ctx->scratch_addr)));
    HIPCHK(hipMemcpyAsync(host_counters, ctx->scratch_addr, sizeof(kpt_trace_counters), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    host_counters->rays = (uint64_t)n;
    return KPT_OK;
}

int kpt_set_settings(kpt_ctx* ctx, const kpt_settings* s) {
    if (!ctx || !s) return KPT_ERR_INVALID_ARG;
    if (s->max_depth < 0 || s->seed_mode < 0 || s->seed_mode > 1 || s->integrator < 0 || s->integrator > 1 ||
        s->postprocess_enable_effects < 0 || s->postprocess_enable_effects > 1 || s->upscale_enable < 0 ||
        s->upscale_enable > 1 || s->upscale_fsr < 0 || s->upscale_fsr > 1)
        return fail(ctx, KPT_ERR_INVALID_ARG, "invalid settings");
    ctx->settings = *s;
    return kpt_reset_accumulation(ctx);   // renderer.cu:301-305
}

int kpt_get_settings(const kpt_ctx* ctx, kpt_settings* out) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    *out = ctx->settings;
    return KPT_OK;
}

int kpt_set_view(kpt_ctx* ctx, const float projection[16], const float view[16]) {
    if (!ctx || !projection || !view) return KPT_ERR_INVALID_ARG;
    hm::Mat4 ip = hm::glm_inverse(hm::load(projection));
    hm::Mat4 iv = hm::glm_inverse(hm::load(view));
    std::memcpy(ctx->cam.prev_inv_view, ctx->cam.inv_view, 64);
    std::memcpy(ctx->cam.prev_inv_proj, ctx->cam.inv_proj, 64);
    std::memcpy(ctx->cam.inv_proj, ip.m, 64);
    std::memcpy(ctx->cam.inv_view, iv.m, 64);
    if (!ctx->settings.use_temporal_accumulation) return kpt_reset_accumulation(ctx);
    return KPT_OK;
}

int kpt_render_views(kpt_ctx* ctx, const kpt_view* views, int32_t n, int32_t spp, float* out) {
    if (!ctx || !views || n <= 0 || spp < 0 || !out) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    if (!ctx->scene_ready) return fail(ctx, KPT_ERR_STATE, "no scene");
    if (n > 65536) return fail(ctx, KPT_ERR_INVALID_ARG, "too many views");
    std::vector<ViewConst> vc((size_t)n);
    int64_t total = 0;
    for (int32_t v = 0; v < n; ++v) {
        const kpt_view& kv = views[v];
        if (kv.width <= 0 || kv.height <= 0) return fail(ctx, KPT_ERR_INVALID_ARG, "view with an empty resolution");
        // Renderer::setView's inverses (renderer.cu:342-349)
        hm::Mat4 ip = hm::glm_inverse(hm::load(kv.projection));
        hm::Mat4 iv = hm::glm_inverse(hm::load(kv.view));
        ViewConst& c = vc[(size_t)v];
        std::memset(&c, 0, sizeof(c));
        std::memcpy(c.inv_view, iv.m, 64);
        std::memcpy(c.inv_proj, ip.m, 64);
        c.cam_pos[0] = iv.m[12]; c.cam_pos[1] = iv.m[13]; c.cam_pos[2] = iv.m[14];
        c.width = kv.width;
        c.height = kv.height;
        c.offset = total;
        total += (int64_t)kv.width * kv.height;
    }
    if (total >= ((int64_t)1 << 31)) return fail(ctx, KPT_ERR_INVALID_ARG, "batch exceeds 2^31 pixels");
    ENSURE_TLAS(ctx);
    int rc = alloc_wavefront(ctx, total);
    if (rc != KPT_OK) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(dev_upload(&ctx->d_views, vc));
    ViewBatch vb;
    vb.views = ctx->d_views;
    vb.nviews = n;
    vb.total = total;
    vb.out = (float4*)out;
    HIPCHK(launch_clear(ctx->stream, vb.out, total));
    SceneView S = scene_view(ctx);
    ctx->ws.heat = nullptr;   // the heat map covers single-view renders only
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    // with the graph knob, each sample's launches are captured and replayed
    // as one graph (the library radix sort stays out of captured graphs)
    const bool graph = ctx->use_graph && !ctx->count_trace;
    TraceTuning tune = ctx->tune;
    if (graph) tune.ray_sort = 0;
    for (int32_t k = 0; k < spp; ++k) {
        FrameConst fc = make_frame(ctx, k);
        auto enqueue = [&]() -> hipError_t {
            return wavefront_frame(ctx->stream, S, fc, k, ctx->fb, ctx->ws, ctx->d_overflow,
                                   (int32_t)ctx->materials.size(), ctx->sort_by_material, tune, nullptr, nullptr,
                                   &vb);
        };
        if (graph) HIPCHK(replay_captured(ctx, &ctx->views_graph_exec, enqueue));
        else HIPCHK(enqueue());
    }
    HIPCHK(launch_views_mean(ctx->stream, vb.out, total, spp));
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->timing[0] = ms;
    ctx->timing[1] = 0.0;
    ctx->timing[2] = 0.0;
    return report_overflow(ctx);
}

int kpt_render_views_host(kpt_ctx* ctx, const kpt_view* views, int32_t n, int32_t spp, float* host_out) {
    if (!ctx || !views || n <= 0 || !host_out) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    int64_t total = 0;
    for (int32_t v = 0; v < n; ++v) {
        if (views[v].width <= 0 || views[v].height <= 0) return fail(ctx, KPT_ERR_INVALID_ARG, "view with an empty resolution");
        total += (int64_t)views[v].width * views[v].height;
    }
    float* d = nullptr;
    HIPCHK(hipMalloc((void**)&d, (size_t)total * 16));
    int rc = kpt_render_views(ctx, views, n, spp, d);
    if (rc == KPT_OK) {
        hipError_t e = hipMemcpy(host_out, d, (size_t)total * 16, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(ctx, KPT_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    (void)hipFree(d);
    return rc;
}

int kpt_set_camera(kpt_ctx* ctx, const kpt_camera* cam) {
    if (!ctx || !cam) return KPT_ERR_INVALID_ARG;
    ctx->cam = *cam;
    return KPT_OK;
}

int kpt_get_camera(const kpt_ctx* ctx, kpt_camera* out) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    *out = ctx->cam;
    return KPT_OK;
}

// Renderer::resizeResolution (renderer.cu:55-101): the output size is
// width x height; the render size is the same, or int(width / 2.5) x
// int(height / 2.5) when upscale_enable is set.  Render-size buffers: render,
// accumulation, G-buffers, V-buffer; output-size: output, back buffer, the
// upscaled render, the bloom mip chain (MipChain::resize, mipchain.cuh:17-35).
// The projection is recomputed for the render size with its field of view
// kept (2 atan(1 / P[1][1])), near 0.1 and far 100, as the reference does on
// every resize (renderer.cu:78-84); a following kpt_set_view replaces it.
int kpt_resize(kpt_ctx* ctx, int32_t width, int32_t height) {
    if (!ctx || width <= 0 || height <= 0 || (int64_t)width * height > (int64_t)1 << 28)
        return KPT_ERR_INVALID_ARG;
    ctx->cache_adapt = false;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    const bool up = ctx->settings.upscale_enable != 0;
    const int32_t rw = up ? (int32_t)((float)width / 2.5f) : width;
    const int32_t rh = up ? (int32_t)((float)height / 2.5f) : height;
    if (rw <= 0 || rh <= 0) return fail(ctx, KPT_ERR_INVALID_ARG, "render size below one pixel");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    free_frame(ctx);
    const size_t n = (size_t)rw * rh, no = (size_t)width * height;
    HIPCHK(hipMalloc((void**)&ctx->fb.accum, n * sizeof(float4)));
    HIPCHK(hipMalloc((void**)&ctx->fb.render, n * sizeof(float4)));
    HIPCHK(hipMalloc((void**)&ctx->fb.gbuffer, n * sizeof(uint4)));
    HIPCHK(hipMalloc((void**)&ctx->fb.prev_gbuffer, n * sizeof(uint4)));
    HIPCHK(hipMalloc((void**)&ctx->fb.vbuffer, n * sizeof(float4)));
    HIPCHK(hipMalloc((void**)&ctx->fb.output, no * sizeof(float4)));
    HIPCHK(hipMalloc((void**)&ctx->pb.back, no * sizeof(float4)));
    HIPCHK(hipMalloc((void**)&ctx->pb.up, no * sizeof(float4)));
    ctx->pb.levels = bloom_levels(width, height);
    for (int k = 0; k <= ctx->pb.levels; ++k) {
        const size_t mn = (size_t)(width >> k) * (size_t)(height >> k);
        HIPCHK(hipMalloc((void**)&ctx->pb.mip[k], mn * sizeof(float4)));
        HIPCHK(launch_clear(ctx->stream, ctx->pb.mip[k], (int64_t)mn));
    }
    ctx->width = rw;
    ctx->height = rh;
    ctx->out_width = width;
    ctx->out_height = height;
    HIPCHK(launch_clear(ctx->stream, (float4*)ctx->fb.prev_gbuffer, (int64_t)n));
    HIPCHK(launch_clear(ctx->stream, ctx->fb.vbuffer, (int64_t)n));
    HIPCHK(launch_clear(ctx->stream, (float4*)ctx->fb.gbuffer, (int64_t)n));
    HIPCHK(launch_clear(ctx->stream, ctx->fb.render, (int64_t)n));
    HIPCHK(launch_clear(ctx->stream, ctx->fb.output, (int64_t)no));
    HIPCHK(launch_clear(ctx->stream, ctx->pb.up, (int64_t)no));
    const hm::Mat4 old_proj = hm::ref_inverse(hm::load(ctx->cam.inv_proj));
    const float fov = 2.0f * std::atan(1.0f / old_proj.m[5]);
    if (std::isfinite(fov) && fov > 0.0f) {
        const hm::Mat4 P = hm::perspective_fov_lh(fov, (float)rw, (float)rh, 0.1f, 100.0f);
        std::memcpy(ctx->cam.inv_proj, hm::ref_inverse(P).m, 64);
    }
    return kpt_reset_accumulation(ctx);
}

int kpt_get_resolution(const kpt_ctx* ctx, int32_t out[4]) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    out[0] = ctx->width;
    out[1] = ctx->height;
    out[2] = ctx->out_width;
    out[3] = ctx->out_height;
    return KPT_OK;
}

int kpt_reset_accumulation(kpt_ctx* ctx) {
    if (!ctx) return KPT_ERR_INVALID_ARG;
    ctx->frame_index = 0;
    if (ctx->host_only || !ctx->fb.accum) return KPT_OK;
    HIPCHK(launch_clear(ctx->stream, ctx->fb.accum, (int64_t)ctx->width * ctx->height));
    return KPT_OK;
}

int kpt_render(kpt_ctx* ctx, int32_t spp) {
    if (!ctx || spp < 0) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    if (!ctx->scene_ready) return fail(ctx, KPT_ERR_STATE, "no scene");
    if (!ctx->fb.accum) return fail(ctx, KPT_ERR_STATE, "no frame buffers (kpt_resize first)");
    ENSURE_TLAS(ctx);
    SceneView S = scene_view(ctx);
    const int64_t npx = (int64_t)ctx->width * ctx->height;
    if (ctx->settings.integrator == KPT_INTEGRATOR_WAVEFRONT) {
        int rc = alloc_wavefront(ctx, npx);
        if (rc != KPT_OK) return rc;
    }
    // radiance cache: wavefront only
    const kpt_cache_settings& cs = ctx->cache_set;
    const bool use_cache = cs.enabled || cs.train;
    if (use_cache && ctx->settings.integrator != KPT_INTEGRATOR_WAVEFRONT)
        return fail(ctx, KPT_ERR_UNSUPPORTED, "the radiance cache runs in the wavefront integrator");
    CacheHooks hooks{};
    const int32_t batch = std::max(1, (cs.train_samples + std::max(1, cs.train_steps) - 1) / std::max(1, cs.train_steps));
    if (use_cache) {
        if (!ctx->cache_weights) return fail(ctx, KPT_ERR_STATE, "radiance cache not configured");
        int rc = cache_alloc_params(ctx, batch);
        // room for the record count's frame-to-frame spread around train_samples
        if (rc == KPT_OK) rc = cache_alloc_records(ctx, npx, std::max(1, 2 * cs.train_samples));
        if (rc != KPT_OK) return rc;
        hooks.cb = &ctx->cb;
        hooks.cf.enabled = cs.enabled;
        hooks.cf.train = cs.train;
        hooks.cf.nee = cs.nee_at_vertex;
        hooks.cf.wide_roughness = cs.wide_roughness;
        const double prob = std::min(1.0, (double)cs.train_samples / (double)std::max<int64_t>(1, npx));
        hooks.cf.train_threshold = (uint32_t)std::min(16777216.0, std::ceil(prob * 16777216.0));
        hooks.cf.train_samples = cs.train_samples;
        hooks.enc = cache_encode(ctx);
        hooks.num_cus = ctx->num_cus;
    }
    if (ctx->count_trace) {
        HIPCHK(hipMemsetAsync(ctx->d_counts, 0, 16 * sizeof(unsigned long long), ctx->stream));
    }
    // per-pixel DebugData of counted wavefront frames (cleared per frame by wavefront_frame)
    ctx->heat_valid = false;
    ctx->ws.heat = nullptr;
    if (ctx->count_trace && ctx->settings.integrator == KPT_INTEGRATOR_WAVEFRONT && spp > 0) {
        if (ctx->heat_pixels != npx) {
            dev_free(&ctx->d_heat);
            ctx->heat_pixels = 0;
            HIPCHK(hipMalloc((void**)&ctx->d_heat, (size_t)npx * 2 * sizeof(int32_t)));
            ctx->heat_pixels = npx;
        }
        ctx->ws.heat = ctx->d_heat;
        ctx->heat_valid = true;
    }
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    double trace_ms = 0.0, tail_ms = 0.0;
    int trace_launches = 0, tail_launches = 0;
    double x_ms[WF_XSTAGES] = {0.0, 0.0, 0.0, 0.0};
    int x_launches[WF_XSTAGES] = {0, 0, 0, 0};
    const bool wavefront = ctx->settings.integrator == KPT_INTEGRATOR_WAVEFRONT;
    const bool graph = ctx->use_graph && !ctx->count_trace;
    // trace times of a frame whose ring slot is about to be reused (or at the end)
    auto harvest = [&](WfInstrument& ins) -> hipError_t {
        hipError_t e = hipSuccess;
        for (int x = 0; e == hipSuccess && x < WF_XSTAGES; ++x) {
            if (ins.xlaunches[x] <= 0) continue;
            float t = 0.0f;
            e = hipEventSynchronize(ins.xend[x]);
            if (e == hipSuccess) e = hipEventElapsedTime(&t, ins.xbegin[x], ins.xend[x]);
            x_ms[x] += t;
            x_launches[x] += ins.xlaunches[x];
            ins.xlaunches[x] = 0;
        }
        if (e != hipSuccess || ins.launches <= 0) return e;
        e = hipEventSynchronize(ins.end[ins.launches - 1]);
        for (int j = 0; e == hipSuccess && j < ins.launches; ++j) {
            float t = 0.0f;
            e = hipEventElapsedTime(&t, ins.begin[j], ins.end[j]);
            trace_ms += t;
        }
        trace_launches += ins.launches;
        ins.launches = 0;
        if (e == hipSuccess && ins.tail_launches > 0) {
            float t = 0.0f;
            e = hipEventSynchronize(ins.tail_end);
            if (e == hipSuccess) e = hipEventElapsedTime(&t, ins.tail_begin, ins.tail_end);
            tail_ms += t;
            tail_launches += ins.tail_launches;
            ins.tail_launches = 0;
        }
        return e;
    };
    for (int r = 0; r < kpt_ctx::INS_RING; ++r) {
        ctx->ins[r].launches = 0;
        ctx->ins[r].tail_launches = 0;
        for (int x = 0; x < WF_XSTAGES; ++x) ctx->ins[r].xlaunches[x] = 0;
    }
    // Two frames in flight: odd frames run their path stages on a second
    // stream with their own path state, so one frame's stages fill the GPU
    // while the other's drain.  Frames stay independent up to their finish:
    // each frame's finishing launch (accumulation, G-buffer) waits for the
    // previous frame's outputs (finish, post chain, G-buffer copy), so the
    // outputs are the sequential frames' bit for bit.  A frame with the
    // radiance cache waits for the previous frame already before its bounce-1
    // work (its cache vertex, queries and records), so only the primary
    // stages overlap and the cache's records, weights and Adam state pass
    // from frame to frame as before.  Counted frames and captured graphs run
    // one at a time.
    const bool overlap = wavefront && !graph && !ctx->count_trace && ctx->overlap && spp > 1;
    // whatever the second stream still holds (a render that stopped on an
    // error) comes before this render's work
    if (ctx->stream2 && ctx->lane_done[0]) {
        HIPCHK(hipEventRecord(ctx->lane_done[0], ctx->stream2));
        HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->lane_done[0], 0));
    }
    if (overlap) {
        if (!ctx->stream2) HIPCHK(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
        for (int k = 0; k < 2; ++k)
            if (!ctx->lane_done[k]) HIPCHK(hipEventCreateWithFlags(&ctx->lane_done[k], hipEventDisableTiming));
        const int rc = alloc_wf_state(ctx, ctx->ws2, npx);
        if (rc != KPT_OK) return rc;
        ctx->ws2.heat = nullptr;
        // the second stream starts after everything enqueued on the first
        HIPCHK(hipEventRecord(ctx->lane_done[1], ctx->stream));
        HIPCHK(hipStreamWaitEvent(ctx->stream2, ctx->lane_done[1], 0));
    }
    hipEvent_t prev_done = nullptr;
    for (int32_t k = 0; k < spp; ++k) {
        const int lane = overlap ? (k & 1) : 0;
        hipStream_t const fs = lane ? ctx->stream2 : ctx->stream;
        WavefrontState& fw = lane ? ctx->ws2 : ctx->ws;
        FrameConst fc = make_frame(ctx, ctx->frame_index);
        WfInstrument* ins = nullptr;
        if (wavefront && !graph) {
            ins = &ctx->ins[k % kpt_ctx::INS_RING];
            HIPCHK(harvest(*ins));
            ins->counters = ctx->count_trace ? ctx->d_counts : nullptr;
        }
        hooks.cf.frame = (uint32_t)ctx->frame_index;
        hooks.cf.adapt = ctx->cache_adapt ? 1 : 0;
        // one executeRendering frame, enqueued without a host synchronisation
        auto enqueue = [&]() -> hipError_t {
            hipError_t e;
            if (wavefront) {
                // the library radix sort stays out of captured graphs
                TraceTuning tune = ctx->tune;
                if (graph) tune.ray_sort = 0;
                e = wavefront_frame(fs, S, fc, ctx->frame_index, ctx->fb, fw, ctx->d_overflow,
                                    (int32_t)ctx->materials.size(), ctx->sort_by_material, tune, ins,
                                    use_cache ? &hooks : nullptr, nullptr, prev_done);
                if (e != hipSuccess) return e;
                if (use_cache && cs.train) {
                    // Adam steps over this frame's suffixes (they shape the next frames)
                    if (ins) (void)hipEventRecord(ins->xbegin[WF_X_TRAIN], fs);
                    for (int32_t st = 0; st < std::max(1, cs.train_steps); ++st) {
                        e = launch_cache_train_step(fs, ctx->cb, hooks.enc, ctx->cb.t_pos, ctx->cb.t_dir,
                                                    ctx->cb.t_target, ctx->cb.counters + 1, 0, ctx->cb.train_cap,
                                                    (int64_t)st * batch, batch, cs.learning_rate,
                                                    ++ctx->cache_steps);
                        if (e != hipSuccess) return e;
                    }
                    if (ins) {
                        (void)hipEventRecord(ins->xend[WF_X_TRAIN], fs);
                        ins->xlaunches[WF_X_TRAIN] = 1;
                    }
                }
            } else {
                e = launch_megakernel(ctx->stream, S, fc, ctx->frame_index, ctx->fb, ctx->d_overflow);
                if (e != hipSuccess) return e;
            }
            // modulate, auto exposure, bloom, post (renderer.cu:117-170)
            PostParams pp{};
            pp.width = ctx->width;
            pp.height = ctx->height;
            pp.temporal = ctx->settings.use_temporal_accumulation ? 1 : 0;
            pp.auto_exposure = ctx->settings.enable_auto_exposure ? 1 : 0;
            pp.bloom = ctx->settings.bloom_generate_bloom ? 1 : 0;
            pp.karis = ctx->settings.bloom_use_karis_average ? 1 : 0;
            pp.bloom_final_blend = ctx->settings.bloom_final_blend;
            pp.bloom_internal_blend = ctx->settings.bloom_internal_blend;
            pp.frame_delta_ms = ctx->frame_delta_ms;
            pp.effects = ctx->settings.postprocess_enable_effects ? 1 : 0;
            pp.out_width = ctx->out_width;
            pp.out_height = ctx->out_height;
            pp.upscale = (ctx->settings.upscale_enable || ctx->out_width != ctx->width ||
                          ctx->out_height != ctx->height) ? 1 : 0;
            pp.fsr = (pp.upscale && ctx->settings.upscale_fsr) ? 1 : 0;
            e = launch_post_frame(fs, ctx->fb, ctx->pb, pp, ctx->num_cus);
            if (e != hipSuccess) return e;
            // end of executeRendering: gbuffer -> prev_gbuffer (renderer.cu:186)
            return hipMemcpyAsync(ctx->fb.prev_gbuffer, ctx->fb.gbuffer, (size_t)npx * sizeof(uint4),
                                  hipMemcpyDeviceToDevice, fs);
        };
        if (graph) {
            // the frame's ~80 launches captured once and replayed as one graph;
            // later frames update the executable graph in place
            HIPCHK(replay_captured(ctx, &ctx->graph_exec, enqueue));
        } else {
            HIPCHK(enqueue());
        }
        if (overlap) {
            HIPCHK(hipEventRecord(ctx->lane_done[lane], fs));
            prev_done = ctx->lane_done[lane];
        }
        if (!wavefront) trace_launches += 1;
        if (use_cache) ctx->cache_adapt = cs.train != 0;
        // camera prev <- curr, frame_index++ (renderer.cu:188-194)
        std::memcpy(ctx->cam.prev_inv_view, ctx->cam.inv_view, 64);
        std::memcpy(ctx->cam.prev_inv_proj, ctx->cam.inv_proj, 64);
        ctx->frame_index++;
    }
    // the first stream takes up what the second one enqueued
    if (overlap) HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->lane_done[1], 0));
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    for (int r = 0; r < kpt_ctx::INS_RING; ++r) HIPCHK(harvest(ctx->ins[r]));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->timing[0] = ms;
    ctx->timing[1] = (wavefront && !graph) ? trace_ms : (wavefront ? 0.0 : ms);
    ctx->timing[2] = trace_launches;
    const bool timed = wavefront && !graph;
    for (int k = 0; k < 6; ++k) ctx->stages[k] = kpt_stage_stats{};
    ctx->stages[0].ms = timed ? trace_ms : 0.0;
    ctx->stages[0].launches = timed ? trace_launches : 0;
    ctx->stages[1].ms = timed ? tail_ms : 0.0;
    ctx->stages[1].launches = timed ? tail_launches : 0;
    for (int x = 0; x < WF_XSTAGES; ++x) {
        ctx->stages[2 + x].ms = timed ? x_ms[x] : 0.0;
        ctx->stages[2 + x].launches = timed ? x_launches[x] : 0;
    }
    if (ctx->count_trace) {
        unsigned long long h[16];
        HIPCHK(hipMemcpy(h, ctx->d_counts, sizeof(h), hipMemcpyDeviceToHost));
        ctx->render_counts.blas_interior_pops = h[0];
        ctx->render_counts.blas_leaf_pops = h[1];
        ctx->render_counts.tri_tests = h[2];
        ctx->render_counts.blas_entries = h[3];
        ctx->render_counts.tlas_pops = h[4];
        ctx->render_counts.rays = 0;
        ctx->stages[0].events = ctx->render_counts;
        kpt_trace_counters& t = ctx->stages[1].events;
        t.blas_interior_pops = h[8];
        t.blas_leaf_pops = h[9];
        t.tri_tests = h[10];
        t.blas_entries = h[11];
        t.tlas_pops = h[12];
        t.rays = 0;
        ctx->stages[1].paths = h[13];
        ctx->stages[1].vertices = h[14];
    }
    return report_overflow(ctx);
}

int kpt_set_exposure(kpt_ctx* ctx, const kpt_exposure* e) {
    if (!ctx || !e) return KPT_ERR_INVALID_ARG;
    // AutoExposureProgram::recordValues + getStandardOutputBasedExposure + Film
    ctx->h_expo.ae = *e;
    const float q = 0.65f;
    const float l_avg = (10.0f / q) * ((e->aperture_f_num * e->aperture_f_num) / (e->iso * e->shutter_speed_secs));
    ctx->h_expo.scalar = 0.18f / l_avg;
    ctx->h_expo.white_ev = e->white_point_ev;
    ctx->h_expo.black_ev = e->black_point_ev;
    if (ctx->host_only) return KPT_OK;
    // the adapted luminance stays as the device left it
    HIPCHK(hipMemcpyAsync(ctx->pb.expo, &ctx->h_expo, offsetof(ExposureState, avg_lum), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

static int read_exposure(kpt_ctx* ctx) {
    if (ctx->host_only) return KPT_OK;
    HIPCHK(hipMemcpyAsync(&ctx->h_expo, ctx->pb.expo, sizeof(ExposureState), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

int kpt_get_exposure(kpt_ctx* ctx, kpt_exposure* out) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    int rc = read_exposure(ctx);
    if (rc != KPT_OK) return rc;
    *out = ctx->h_expo.ae;
    return KPT_OK;
}

int kpt_get_film(kpt_ctx* ctx, kpt_film_state* out) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    int rc = read_exposure(ctx);
    if (rc != KPT_OK) return rc;
    out->luminance_exposure_scalar = ctx->h_expo.scalar;
    out->white_point_ev = ctx->h_expo.white_ev;
    out->black_point_ev = ctx->h_expo.black_ev;
    out->scene_average_luminance = ctx->h_expo.avg_lum;
    return KPT_OK;
}

int kpt_set_frame_delta(kpt_ctx* ctx, float delta_time_ms) {
    if (!ctx || !(delta_time_ms >= 0.0f)) return KPT_ERR_INVALID_ARG;
    ctx->frame_delta_ms = delta_time_ms;
    return KPT_OK;
}

// ---------------------------------------------------------------- radiance cache
int kpt_cache_configure(kpt_ctx* ctx, const kpt_cache_settings* s) {
    if (!ctx || !s) return KPT_ERR_INVALID_ARG;
    ctx->cache_adapt = false;
    if (s->train_samples < 0 || s->train_samples > (1 << 22) || s->train_steps < 0 || s->train_steps > 64 ||
        !(s->learning_rate >= 0.0f) || !(s->radiance_scale > 0.0f) || (s->nee_at_vertex & ~1) ||
        !(s->wide_roughness >= 0.0f))
        return fail(ctx, KPT_ERR_INVALID_ARG, "cache settings out of range");
    if ((s->enabled || s->train) && s->train_samples > 0 && s->train_steps == 0)
        return fail(ctx, KPT_ERR_INVALID_ARG, "training needs at least one step per frame");
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    const bool reseed = !ctx->cache_weights || s->seed != ctx->cache_set.seed;
    ctx->cache_set = *s;
    const int32_t steps = std::max(1, s->train_steps);
    int rc = cache_alloc_params(ctx, (std::max(1, s->train_samples) + steps - 1) / steps);
    if (rc != KPT_OK) return rc;
    if (reseed) {
        std::vector<float> w = cache_init_weights(s->seed);
        rc = cache_upload_weights(ctx, w.data());
        if (rc != KPT_OK) return rc;
    }
    return KPT_OK;
}

int kpt_cache_get_settings(const kpt_ctx* ctx, kpt_cache_settings* out) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    *out = ctx->cache_set;
    return KPT_OK;
}

int kpt_cache_get_weights(kpt_ctx* ctx, float* host_out, int64_t n) {
    if (!ctx || !host_out || n < RC_PARAMS) return KPT_ERR_INVALID_ARG;
    if (!ctx->cache_weights) return fail(ctx, KPT_ERR_STATE, "radiance cache not configured");
    HIPCHK(hipMemcpyAsync(host_out, ctx->cb.w, RC_PARAMS * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

int kpt_cache_set_weights(kpt_ctx* ctx, const float* host_w, int64_t n) {
    if (!ctx || !host_w || n < RC_PARAMS) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    int rc = cache_alloc_params(ctx, 1);
    if (rc != KPT_OK) return rc;
    return cache_upload_weights(ctx, host_w);
}

int kpt_cache_infer(kpt_ctx* ctx, const float* positions, const float* directions, int64_t n, float* out) {
    if (!ctx || n < 0 || (n > 0 && (!positions || !directions || !out))) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    if (!ctx->cache_weights) return fail(ctx, KPT_ERR_STATE, "radiance cache not configured");
    HIPCHK(launch_cache_infer_raw(ctx->stream, ctx->cb, cache_encode(ctx), (const float4*)positions,
                                  (const float4*)directions, n, (float4*)out, ctx->num_cus));
    return KPT_OK;
}

int kpt_cache_train(kpt_ctx* ctx, const float* positions, const float* directions, const float* targets, int64_t n,
                    int32_t steps, float* host_loss) {
    if (!ctx || n < 0 || steps <= 0 || (n > 0 && (!positions || !directions || !targets)))
        return KPT_ERR_INVALID_ARG;
    if (n > ((int64_t)1 << 26)) return fail(ctx, KPT_ERR_INVALID_ARG, "too many training samples");
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    if (!ctx->cache_weights) return fail(ctx, KPT_ERR_STATE, "radiance cache not configured");
    const int32_t batch = (int32_t)std::max<int64_t>(1, (n + steps - 1) / steps);
    int rc = cache_alloc_params(ctx, batch);
    if (rc != KPT_OK) return rc;
    const CacheEncode enc = cache_encode(ctx);
    for (int32_t st = 0; st < steps; ++st) {
        HIPCHK(launch_cache_train_step(ctx->stream, ctx->cb, enc, (const float4*)positions, (const float4*)directions,
                                       (const float4*)targets, nullptr, n, n, (int64_t)st * batch, batch,
                                       ctx->cache_set.learning_rate, ++ctx->cache_steps));
    }
    if (host_loss) {
        HIPCHK(hipMemcpyAsync(host_loss, ctx->cb.stats, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return KPT_OK;
}

int kpt_cache_get_gradient(kpt_ctx* ctx, float* host_out, int64_t n) {
    if (!ctx || !host_out || n < RC_PARAMS) return KPT_ERR_INVALID_ARG;
    if (!ctx->cache_weights) return fail(ctx, KPT_ERR_STATE, "radiance cache not configured");
    HIPCHK(hipMemcpyAsync(host_out, ctx->cb.grad, RC_PARAMS * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

int kpt_cache_get_queries(kpt_ctx* ctx, float* pos, float* dir, float* beta, int64_t cap, int64_t* count) {
    if (!ctx || !count || cap < 0 || (cap > 0 && (!pos || !dir || !beta))) return KPT_ERR_INVALID_ARG;
    *count = 0;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    if (!ctx->cb.counters || !ctx->cb.q_pos) return KPT_OK;
    int32_t c = 0;
    HIPCHK(hipMemcpyAsync(&c, ctx->cb.counters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int64_t n = std::min<int64_t>(std::min<int64_t>(c, ctx->cb.query_cap), cap);
    *count = std::min<int64_t>(c, ctx->cb.query_cap);
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(pos, ctx->cb.q_pos, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(dir, ctx->cb.q_dir, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(beta, ctx->cb.q_beta, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return KPT_OK;
}

int kpt_cache_stats(kpt_ctx* ctx, double out[5]) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    for (int k = 0; k < 5; ++k) out[k] = 0.0;
    out[4] = (double)ctx->cache_steps;
    if (ctx->host_only || !ctx->cache_weights) return KPT_OK;
    float st[2];
    int32_t cnt[2];
    HIPCHK(hipMemcpyAsync(st, ctx->cb.stats, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(cnt, ctx->cb.counters, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    out[0] = st[0];
    out[1] = st[1];
    out[2] = cnt[0];
    out[3] = cnt[1];
    return KPT_OK;
}

int kpt_get_frame_index(const kpt_ctx* ctx, int32_t* out) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    *out = ctx->frame_index;
    return KPT_OK;
}

int kpt_readback(kpt_ctx* ctx, int32_t which, void* host_dst, int64_t bytes) {
    if (!ctx || !host_dst) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    const void* src = nullptr;
    switch (which) {
        case 0: src = ctx->fb.render; break;
        case 1: src = ctx->fb.accum; break;
        case 2: src = ctx->fb.gbuffer; break;
        case 3: src = ctx->fb.vbuffer; break;
        case 4: src = ctx->fb.output; break;
        default: return fail(ctx, KPT_ERR_INVALID_ARG, "unknown buffer");
    }
    if (!src) return fail(ctx, KPT_ERR_STATE, "no frame buffers");
    const int64_t need = (which == 4 ? (int64_t)ctx->out_width * ctx->out_height
                                     : (int64_t)ctx->width * ctx->height) * 16;
    if (bytes < need) return fail(ctx, KPT_ERR_INVALID_ARG, "destination too small");
    HIPCHK(hipMemcpyAsync(host_dst, src, (size_t)need, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

int kpt_sync(kpt_ctx* ctx) {
    if (!ctx) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return KPT_OK;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->stream2) HIPCHK(hipStreamSynchronize(ctx->stream2));
    return report_overflow(ctx);
}

int kpt_set_tuning(kpt_ctx* ctx, const int32_t* knobs, int32_t n) {
    if (!ctx || !knobs || n < 5) return KPT_ERR_INVALID_ARG;
    if (knobs[0] < 0 || knobs[0] > 1 || knobs[1] < 64 || knobs[1] > 65536 || knobs[2] < 1 || knobs[2] > 64 ||
        knobs[3] < 1 || knobs[3] > 64 || knobs[4] < 0 || knobs[4] > 2)
        return fail(ctx, KPT_ERR_INVALID_ARG, "tuning knob out of range");
    ctx->tune.refill = knobs[0];
    ctx->tune.chunk = knobs[1];
    ctx->tune.refill_min = knobs[2];
    ctx->tune.blocks_per_cu = knobs[3];
    ctx->sort_by_material = knobs[4];
    if (n >= 6) {
        if (knobs[5] < 0 || knobs[5] > 1) return fail(ctx, KPT_ERR_INVALID_ARG, "tuning knob out of range");
        ctx->use_graph = knobs[5];
    }
    if (n >= 7) {
        if (knobs[6] < 0 || knobs[6] > 1) return fail(ctx, KPT_ERR_INVALID_ARG, "tuning knob out of range");
        ctx->tune.ray_sort = knobs[6];
    }
    if (n >= 8) {
        if (knobs[7] < 0 || knobs[7] > 64) return fail(ctx, KPT_ERR_INVALID_ARG, "tuning knob out of range");
        ctx->tune.tail = knobs[7];
    }
    if (n >= 9) {
        if (knobs[8] < 0 || knobs[8] > 2) return fail(ctx, KPT_ERR_INVALID_ARG, "tuning knob out of range");
        if (knobs[8] != ctx->tlas_device && ctx->scene_ready) ctx->tlas_dirty = true;   // rebuilt for the new mode
        ctx->tlas_device = knobs[8];
    }
    if (n >= 10) {
        if (knobs[9] < 0 || knobs[9] > 1) return fail(ctx, KPT_ERR_INVALID_ARG, "tuning knob out of range");
        ctx->blas_device = knobs[9];
    }
    if (n >= 11) {
        if (knobs[10] < LDS_STACK_MIN || knobs[10] > LDS_STACK_MAX)
            return fail(ctx, KPT_ERR_INVALID_ARG, "tuning knob out of range");
        ctx->tune.stack = knobs[10];
    }
    if (n >= 12) {
        if (knobs[11] != 64 && knobs[11] != 128 && knobs[11] != 256)
            return fail(ctx, KPT_ERR_INVALID_ARG, "tuning knob out of range");
        ctx->tune.block = knobs[11];
    }
    if (n >= 14) {
        if (knobs[13] < 0 || knobs[13] > 1) return fail(ctx, KPT_ERR_INVALID_ARG, "tuning knob out of range");
        ctx->overlap = knobs[13];
    }
    if (n >= 13) {
        if (knobs[12] < 0 || knobs[12] > 64) return fail(ctx, KPT_ERR_INVALID_ARG, "tuning knob out of range");
        ctx->tune.suspend = knobs[12];
    }
    return KPT_OK;
}

int kpt_get_tuning(const kpt_ctx* ctx, int32_t* knobs, int32_t n) {
    if (!ctx || !knobs || n < 5) return KPT_ERR_INVALID_ARG;
    knobs[0] = ctx->tune.refill;
    knobs[1] = ctx->tune.chunk;
    knobs[2] = ctx->tune.refill_min;
    knobs[3] = ctx->tune.blocks_per_cu;
    knobs[4] = ctx->sort_by_material;
    if (n >= 6) knobs[5] = ctx->use_graph;
    if (n >= 7) knobs[6] = ctx->tune.ray_sort;
    if (n >= 8) knobs[7] = ctx->tune.tail;
    if (n >= 9) knobs[8] = ctx->tlas_device;
    if (n >= 10) knobs[9] = ctx->blas_device;
    if (n >= 11) knobs[10] = ctx->tune.stack;
    if (n >= 12) knobs[11] = ctx->tune.block;
    if (n >= 13) knobs[12] = ctx->tune.suspend;
    if (n >= 14) knobs[13] = ctx->overlap;
    return KPT_OK;
}

int kpt_set_trace_counting(kpt_ctx* ctx, int32_t enable) {
    if (!ctx) return KPT_ERR_INVALID_ARG;
    ctx->count_trace = enable != 0;
    return KPT_OK;
}

int kpt_get_render_counters(const kpt_ctx* ctx, kpt_trace_counters* out) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    *out = ctx->render_counts;
    return KPT_OK;
}

int kpt_get_heat_map(kpt_ctx* ctx, int32_t* counts, float* rgba, int64_t n_pixels) {
    if (!ctx || n_pixels < 0 || (!counts && !rgba)) return KPT_ERR_INVALID_ARG;
    if (ctx->host_only) return fail(ctx, KPT_ERR_NO_DEVICE, "host-only context");
    if (!ctx->heat_valid || !ctx->d_heat)
        return fail(ctx, KPT_ERR_STATE, "no counted wavefront render (kpt_set_trace_counting before kpt_render)");
    if (n_pixels != ctx->heat_pixels)
        return fail(ctx, KPT_ERR_INVALID_ARG, "n_pixels is not the counted render's width x height");
    if (counts)
        HIPCHK(hipMemcpyAsync(counts, ctx->d_heat, (size_t)n_pixels * 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
    if (rgba) {
        float4* d = nullptr;
        HIPCHK(hipMalloc((void**)&d, (size_t)std::max<int64_t>(1, n_pixels) * sizeof(float4)));
        hipError_t e = launch_heat_colour(ctx->stream, ctx->d_heat, n_pixels, d);
        if (e == hipSuccess)
            e = hipMemcpyAsync(rgba, d, (size_t)n_pixels * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        (void)hipFree(d);
        HIPCHK(e);
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return KPT_OK;
}

int kpt_get_stage_stats(const kpt_ctx* ctx, int32_t stage, kpt_stage_stats* out) {
    if (!ctx || !out || stage < 0 || stage > 5) return KPT_ERR_INVALID_ARG;
    *out = ctx->stages[stage];
    return KPT_OK;
}

int kpt_last_timing(const kpt_ctx* ctx, double out[4]) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    for (int i = 0; i < 4; ++i) out[i] = ctx->timing[i];
    return KPT_OK;
}

int kpt_measure_gather16(const void* src, int64_t src_records, const int32_t* idx, int64_t n, void* dst) {
    if (!src || !idx || !dst || src_records <= 0 || n < 0) return KPT_ERR_INVALID_ARG;
    if (n == 0) return KPT_OK;
    hipLaunchKernelGGL(k_gather16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr,
                       static_cast<const float4*>(src), src_records, idx, n, static_cast<float4*>(dst));
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return KPT_ERR_HIP;
    return KPT_OK;
}

int kpt_last_build_timing(const kpt_ctx* ctx, double out[2]) {
    if (!ctx || !out) return KPT_ERR_INVALID_ARG;
    out[0] = ctx->blas_build_ms[0];
    out[1] = ctx->blas_build_ms[1];
    return KPT_OK;
}

}  // extern "C"
