// kpt_trace.h — two-level BVH2 traversal for gfx950 (device code).
//
// Semantics are those of TLAS::intersect / TLAS::intersectP (tlas.cu:8-125)
// with BLAS2::intersect / intersectP (blas.cuh:50-188), restructured for the
// GPU:
//   * one loop and one stack for both levels; a stack entry carries a level
//     bit, so the BLAS of a TLAS leaf is traversed to completion before the
//     next TLAS entry is popped (LIFO), exactly as the nested reference calls;
//   * the nearer child is kept in a register ("current entry") instead of
//     being pushed and popped at once, and only the farther sibling goes on
//     the stack; the visiting order is the reference's (it pushes the farther
//     child first and pops the nearer one next);
//   * the per-lane stack lives in LDS ([entry][lane] layout: lanes of a wave
//     hit distinct banks) with a private-memory (scratch) spill region for
//     entries beyond the LDS depth, instead of the reference's fixed 16-entry
//     register array that overflows on deep trees;
//   * the BLAS is entered with tmax = min(tmax, closest-so-far).  The reference
//     starts every BLAS with the original tmax and a local closest of +inf;
//     nodes this extra bound skips have a slab entry at or beyond the
//     closest hit, so in exact arithmetic every triangle under them is too,
//     and TLAS::intersect keeps a hit only if strictly closer; the nodes that
//     remain are visited in the reference's order, so the closest hit
//     (including which of two equal-distance hits wins) is the same.  In
//     float the argument has one gap: a triangle whose rounded distance
//     falls below the rounded slab entry of its own box, with another
//     instance's hit in between.  Overlapping coplanar instances a few ulps
//     apart (flat boxes, ties on most rays) never showed it: the oracle's
//     bounded mode equals its reference mode on 600,000 such rays, and these
//     kernels equal the reference mode on the same rays;
//   * any-hit needs no ordering: its result is "some reachable triangle is
//     hit", which does not depend on the visiting order; it visits the
//     nearer child first, which finds occluders close to the origin sooner;
//   * a node reached from its parent reuses the header words read in the
//     parent's child test, and a leaf's triangles are loaded TRI_BATCH at a
//     time (tested in slot order, so the closest-hit tie rule is unchanged).
// Traversal is a state machine (init / step) so persistent kernels can refill
// a lane with a new ray as soon as its ray finishes.
#pragma once
#include "kpt_vec.h"
#include "../common/kpt_device_types.h"

namespace kpt {

constexpr int TRACE_BLOCK = 256;      // threads per workgroup of the fixed-shape trace kernels
constexpr int LDS_STACK = 16;         // stack entries per lane held in LDS (fixed-shape kernels)
constexpr int STACK_TOTAL = 128;      // entries per lane, LDS and private together; a push beyond overflows
// The wavefront's persistent kernels take their LDS depth and workgroup size
// from the tuning knobs (dynamic LDS of depth x workgroup words); the private
// spill array is sized for the shallowest LDS part, so the total capacity is
// STACK_TOTAL whatever the depth.
constexpr int LDS_STACK_MIN = 8;
constexpr int LDS_STACK_MAX = 32;
constexpr int SPILL_STACK = STACK_TOTAL - LDS_STACK_MIN;
constexpr uint32_t TLAS_BIT = 0x80000000u;
#ifndef KPT_TRI_BATCH
#define KPT_TRI_BATCH 4
#endif
constexpr int TRI_BATCH = KPT_TRI_BATCH;   // leaf triangles loaded together

struct Ray {
    v3 o, d, inv;
};
KD Ray make_ray(v3 o, v3 d) { Ray r; r.o = o; r.d = d; r.inv = 1.0f / d; return r; }

struct Hit {
    float t;
    int32_t slot;   // triangle slot (-1 = miss)
    int32_t inst;
    float u, v;
};

struct TraceCounts {
    uint32_t interior, leaf, tri, entries, tlas;
};

// Bounds3f::intersectP, bounds.cuh:38-57
KD bool slab(const float4& lo, const float4& hi, const Ray& r, float tmin, float tmax, float* tenter) {
    float t0x = (lo.x - r.o.x) * r.inv.x, t0y = (lo.y - r.o.y) * r.inv.y, t0z = (lo.z - r.o.z) * r.inv.z;
    float t1x = (hi.x - r.o.x) * r.inv.x, t1y = (hi.y - r.o.y) * r.inv.y, t1z = (hi.z - r.o.z) * r.inv.z;
    float ex = fminf_(t0x, t1x), ey = fminf_(t0y, t1y), ez = fminf_(t0z, t1z);
    float xx = fmaxf_(t0x, t1x), xy = fmaxf_(t0y, t1y), xz = fmaxf_(t0z, t1z);
    float te = fmaxf_(fmaxf_(ex, ey), ez);
    float tx = fminf_(fminf_(xx, xy), xz);
    *tenter = te;
    return (te <= tx) && (tx >= tmin) && (te <= tmax);
}

KD void load_node(const Node* nodes, uint32_t i, float4* lo, float4* hi) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    *lo = p[0];
    *hi = p[1];
}

// Object-space ray: Ray::transform (ray.cuh:25-29) with the full 4-term rows.
KD Ray object_ray(const InstTrav& it, const Ray& w) {
    v3 o = mk3(it.r0[0] * w.o.x + it.r0[1] * w.o.y + it.r0[2] * w.o.z + it.r0[3] * 1.0f,
               it.r1[0] * w.o.x + it.r1[1] * w.o.y + it.r1[2] * w.o.z + it.r1[3] * 1.0f,
               it.r2[0] * w.o.x + it.r2[1] * w.o.y + it.r2[2] * w.o.z + it.r2[3] * 1.0f);
    v3 d = mk3(it.r0[0] * w.d.x + it.r0[1] * w.d.y + it.r0[2] * w.d.z + it.r0[3] * 0.0f,
               it.r1[0] * w.d.x + it.r1[1] * w.d.y + it.r1[2] * w.d.z + it.r1[3] * 0.0f,
               it.r2[0] * w.d.x + it.r2[1] * w.d.y + it.r2[2] * w.d.z + it.r2[3] * 0.0f);
    return make_ray(o, d);
}

// Triangle::intersect (triangle.cuh:66-102) on a slot's precomputed edges.
KD bool tri_test(const TriPos& tp, const Ray& r, float tmin, float tmax, float* t_out, float* u_out, float* v_out) {
    const float EPS = 1.0e-5f;
    v3 v0 = mk3(tp.v0[0], tp.v0[1], tp.v0[2]);
    v3 e1 = mk3(tp.e1[0], tp.e1[1], tp.e1[2]);
    v3 e2 = mk3(tp.e2[0], tp.e2[1], tp.e2[2]);
    v3 pvec = cross(r.d, e2);
    float det = dot(e1, pvec);
    if (det > -EPS && det < EPS) return false;
    float inv = 1.0f / det;
    v3 tvec = r.o - v0;
    float u = inv * dot(tvec, pvec);
    if (u < 0.0f || u > 1.0f) return false;
    v3 q = cross(tvec, e1);
    float v = inv * dot(r.d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = inv * dot(e2, q);
    if (t > EPS && t < tmax && t > tmin) {
        *t_out = t; *u_out = u; *v_out = v;
        return true;
    }
    return false;
}

// Per-lane stack: LDS part [depth][stride] (stride = the workgroup's lanes,
// so the lanes of a wave hit distinct banks) + a private spill array owned by
// the caller (kept outside the traversal state so the state itself stays in
// registers).
typedef __attribute__((address_space(3))) uint32_t lds_u32;   // a pointer that is known to be LDS

// Where a lane's stack lives: its first LDS word, the word stride between
// entries and the number of LDS entries (LDS_STACK_MIN..LDS_STACK_MAX, or
// LDS_STACK for the fixed-shape kernels); all but the pointer are uniform.
struct StackShape {
    uint32_t* lds;
    int stride;
    int depth;
};

struct LaneStack {
    lds_u32* lds;       // the lane's entry 0
    uint32_t* spill;    // private array of SPILL_STACK entries
    int stride, depth;
    int sp;
    bool overflow;
    KD void push(uint32_t e) {
        if (sp < depth) lds[sp * stride] = e;
        else if (sp < STACK_TOTAL) spill[sp - depth] = e;
        else { overflow = true; return; }
        ++sp;
    }
    KD uint32_t pop() {
        --sp;
        return sp < depth ? lds[sp * stride] : spill[sp - depth];
    }
    // entry i (< STACK_TOTAL) read or written in place (a parked trace's copy)
    KD uint32_t at(int i) const { return i < depth ? lds[i * stride] : spill[i - depth]; }
    KD void set(int i, uint32_t e) {
        if (i < depth) lds[i * stride] = e;
        else spill[i - depth] = e;
    }
    KD void reset(const StackShape& sh, uint32_t* spill_) {
        // stride and depth are the same for every lane: read them as such, so
        // they stay in scalar registers wherever the traversal runs
        lds = (lds_u32*)sh.lds; spill = spill_;
        stride = __builtin_amdgcn_readfirstlane(sh.stride);
        depth = __builtin_amdgcn_readfirstlane(sh.depth);
        sp = 0; overflow = false;
    }
};

// ---- stack / current entries.  A BVH2 node is represented by what the
// traversal needs to process it without loading the node itself:
//   interior BLAS node:  pair index of its children (children at 2q, 2q+1)
//   interior TLAS node:  TLAS_BIT | pair index of its children
//   BLAS leaf:           LEAF_BIT | count << 23 | first slot   (count < 64, first < 2^23)
//   larger BLAS leaf:    BIG_BIT | node index (its header is loaded)
//   TLAS leaf:           TLAS_BIT | LEAF_BIT | instance
// Processing an interior entry loads the 192-byte NodeBlock of its children
// pair: both children and their children.  That replays two consecutive steps
// of the BVH2 traversal (the node, then its nearer child when that child is
// interior) from one block load, at the same closest distance h.t — no
// triangle is tested between the two — so every culling decision, every push
// and the visiting order are those of the node-at-a-time traversal.
constexpr uint32_t LEAF_BIT = 0x40000000u;
constexpr uint32_t BIG_BIT = 0x20000000u;
constexpr uint32_t ENTRY_PAYLOAD = 0x1FFFFFFFu;

KD uint32_t bits_u(float f) { return __builtin_bit_cast(uint32_t, f); }

// Entry of a BLAS child given its node index and header words.
KD uint32_t blas_entry(uint32_t idx, uint32_t a, uint32_t b) {
    if (b == 0u) return a >> 1;
    if (a < (1u << 23) && b < 64u) return LEAF_BIT | (b << 23) | a;
    return BIG_BIT | idx;
}
// Entry of a TLAS child given its header words.
KD uint32_t tlas_entry(uint32_t a, uint32_t b) {
    return a != 0u ? (TLAS_BIT | (a >> 1)) : (TLAS_BIT | LEAF_BIT | b);
}

struct BlockView {
    float4 c0l, c0h, c1l, c1h;   // children of the pair's parent
    float4 g0l, g0h, g1l, g1h;   // children of child 0 (when interior)
    float4 g2l, g2h, g3l, g3h;   // children of child 1 (when interior)
};
KD void load_block(const NodeBlock* blocks, uint32_t q, BlockView* v) {
    const float4* p = reinterpret_cast<const float4*>(blocks + q);
    v->c0l = p[0]; v->c0h = p[1]; v->c1l = p[2]; v->c1h = p[3];
    v->g0l = p[4]; v->g0h = p[5]; v->g1l = p[6]; v->g1h = p[7];
    v->g2l = p[8]; v->g2h = p[9]; v->g3l = p[10]; v->g3h = p[11];
}

// What one traversal step of a lane reads, chosen before any of it is loaded
// so that every lane of a wave issues its loads together and the wave waits
// for one memory round trip per step whatever mix of work its lanes hold:
//   STEP_INTERIOR  the 192-byte NodeBlock of the entry's children pair
//   STEP_INSTANCE  the 64-byte InstTrav of a TLAS leaf
//   STEP_TRIS      up to TRI_BATCH 48-byte TriPos of the current leaf
//   STEP_HEADER    the 32-byte node of a leaf too large for an entry
enum : int { STEP_INTERIOR = 0, STEP_INSTANCE = 1, STEP_TRIS = 2, STEP_HEADER = 3 };
constexpr int STEP_LOADS = 12;   // float4 loads per step (192 bytes)
static_assert(3 * TRI_BATCH <= STEP_LOADS, "a triangle batch fits one step's loads");

// Loads float4 [0, n) of base into v (n >= 1); the slots past n repeat the
// last one, so the loads are issued unconditionally and stay in bounds.
KD void step_load(const float4* base, int n, float4 v[STEP_LOADS]) {
#pragma unroll
    for (int k = 0; k < STEP_LOADS; ++k) v[k] = base[k < n ? k : n - 1];
}

// Classifies the lane's next step and points base/n at what it reads.  A
// LEAF entry is decoded here (first slot and count, no load), so its first
// triangles are read in the same step.
KD int step_kind(const SceneView& S, uint32_t cur, uint32_t& la, uint32_t& lb, const float4** base, int* n) {
    if (lb == 0u) {
        if (!(cur & (LEAF_BIT | BIG_BIT))) {
            const NodeBlock* blocks = (cur & TLAS_BIT) ? S.tlas_blocks : S.blas_blocks;
            *base = reinterpret_cast<const float4*>(blocks + (cur & ENTRY_PAYLOAD));
            *n = 12;
            return STEP_INTERIOR;
        }
        if (cur & TLAS_BIT) {
            *base = reinterpret_cast<const float4*>(S.inst_trav + (cur & ENTRY_PAYLOAD));
            *n = 4;
            return STEP_INSTANCE;
        }
        if (!(cur & LEAF_BIT)) {
            *base = reinterpret_cast<const float4*>(S.blas_nodes + (cur & ENTRY_PAYLOAD));
            *n = 2;
            return STEP_HEADER;
        }
        la = cur & 0x7FFFFFu;
        lb = (cur >> 23) & 0x3Fu;
    }
    *base = reinterpret_cast<const float4*>(S.tris + la);
    *n = 3 * (int)(lb < (uint32_t)TRI_BATCH ? lb : (uint32_t)TRI_BATCH);
    return STEP_TRIS;
}

// Closest-hit traversal state of one lane.
struct ClosestTrav {
    Ray w, o;
    float tmin, tmax;
    Hit h;
    int32_t cur_inst;
    uint32_t cur;
    uint32_t la, lb;   // leaf in progress: next triangle slot, triangles left (0: none)
    LaneStack st;

    // TLAS::intersect entry: empty TLAS or a miss of the root box ends the
    // query at once (tlas.cu:62-69).  Returns false when nothing remains.
    KD bool init(const SceneView& S, const Ray& ray, float t0, float t1, const StackShape& sh, uint32_t* spill) {
        w = ray; o = ray; tmin = t0; tmax = t1;
        h.t = __builtin_inff(); h.slot = -1; h.inst = -1; h.u = 0.0f; h.v = 0.0f;
        cur_inst = -1;
        la = 0u; lb = 0u;
        st.reset(sh, spill);
        if (S.tlas_root < 0) return false;
        float4 rlo, rhi;
        load_node(S.tlas_nodes, 0, &rlo, &rhi);
        float te;
        if (!slab(rlo, rhi, w, tmin, tmax, &te)) return false;
        cur = tlas_entry(bits_u(rlo.w), bits_u(rhi.w));
        return true;
    }

    // Continues with the stack top; false when the stack is empty.
    KD bool next() {
        if (st.sp == 0 || st.overflow) return false;
        cur = st.pop();
        return true;
    }

    // One BVH2 interior step on the two children (l1,h1 = index i1; l2,h2 =
    // index i2) of the node being processed: TLAS (level bit set) or BLAS.
    // Returns 0 = nothing hit (pop), 1 = the nearer (or only) child is in
    // *near_* (after the farther one was pushed).
    // One code path for both levels (the level is a per-lane select), so a
    // wave whose lanes sit at different levels does not run it twice.
    KD bool children(bool tl, const float4& l1, const float4& h1, const float4& l2, const float4& h2, uint32_t i1,
                     uint32_t i2, int* near_k) {
        Ray r;
        r.o = tl ? w.o : o.o; r.d = tl ? w.d : o.d; r.inv = tl ? w.inv : o.inv;
        const float bt = tl ? tmax : fminf_(tmax, h.t);
        float c1, c2;
        const bool k1 = slab(l1, h1, r, tmin, bt, &c1) && c1 < h.t;
        const bool k2 = slab(l2, h2, r, tmin, bt, &c2) && c2 < h.t;
        if (k1 && k2) {
            const bool far1 = c1 > c2;
            const uint32_t fa = bits_u(far1 ? l1.w : l2.w), fb = bits_u(far1 ? h1.w : h2.w);
            st.push(tl ? tlas_entry(fa, fb) : blas_entry(far1 ? i1 : i2, fa, fb));
            *near_k = far1 ? 1 : 0;
            return true;
        }
        if (k1) { *near_k = 0; return true; }
        if (k2) { *near_k = 1; return true; }
        return false;
    }

    // Processes the current entry — an interior step, an instance entry, a
    // large leaf's header, or one batch of TRI_BATCH triangles of a leaf (a
    // leaf spans several steps, so a wave's lanes at interior nodes are not
    // held up by another lane's whole leaf); returns false when the query is
    // finished.  Whatever the lane's kind of step, its reads are issued by
    // the same instructions (step_load).
    template <bool COUNT>
    KD bool step(const SceneView& S, TraceCounts* cnt) {
        const uint32_t e = cur;
        const bool new_leaf = (lb == 0u) && (e & LEAF_BIT) && !(e & TLAS_BIT);
        const float4* base;
        int n;
        const int kind = step_kind(S, e, la, lb, &base, &n);
        float4 v[STEP_LOADS];
        step_load(base, n, v);
        if (kind == STEP_INTERIOR) {
            // interior node (TLAS or BLAS): its children pair's block
            const bool tl = (e & TLAS_BIT) != 0;
            const uint32_t q = e & ENTRY_PAYLOAD;
            int nk;
            bool any;
            if (COUNT) { if (tl) cnt->tlas++; else cnt->interior++; }
            any = children(tl, v[0], v[1], v[2], v[3], 2 * q, 2 * q + 1, &nk);
            if (!any) return next();
            if (st.overflow) return false;
            const float4 nl = nk ? v[2] : v[0], nh = nk ? v[3] : v[1];
            const uint32_t na = bits_u(nl.w), nb = bits_u(nh.w);
            const bool n_interior = tl ? (na != 0u) : (nb == 0u);
            if (!n_interior) {
                cur = tl ? tlas_entry(na, nb) : blas_entry(2 * q + (uint32_t)nk, na, nb);
                return true;
            }
            // the nearer child is interior: its step, from the same block
            const float4 al = nk ? v[8] : v[4], ah = nk ? v[9] : v[5];
            const float4 bl = nk ? v[10] : v[6], bh = nk ? v[11] : v[7];
            const uint32_t p = na >> 1;   // the nearer child's children pair
            if (COUNT) { if (tl) cnt->tlas++; else cnt->interior++; }
            any = children(tl, al, ah, bl, bh, 2 * p, 2 * p + 1, &nk);
            if (!any) return next();
            if (st.overflow) return false;
            const float4 ml = nk ? bl : al, mh = nk ? bh : ah;
            cur = tl ? tlas_entry(bits_u(ml.w), bits_u(mh.w)) : blas_entry(2 * p + (uint32_t)nk, bits_u(ml.w), bits_u(mh.w));
            return true;
        }
        if (kind == STEP_INSTANCE) {
            // TLAS leaf: enter the instance's BLAS
            if (COUNT) { cnt->tlas++; cnt->entries++; }
            cur_inst = (int32_t)(e & ENTRY_PAYLOAD);
            InstTrav it;
            __builtin_memcpy(&it, v, sizeof(InstTrav));
            o = object_ray(it, w);
            if (it.root_entry != 0xFFFFFFFFu) { cur = it.root_entry; return true; }
            return next();
        }
        if (kind == STEP_HEADER) {
            // a leaf too large for an entry: its header, the triangles from the next step
            la = bits_u(v[0].w);
            lb = bits_u(v[1].w);
            if (COUNT) cnt->leaf++;
            return true;
        }
        if (COUNT && new_leaf) cnt->leaf++;
        // the next TRI_BATCH triangles of the leaf, tested in slot order
#pragma unroll
        for (int j = 0; j < TRI_BATCH; ++j) {
            if ((uint32_t)j < lb) {
                if (COUNT) cnt->tri++;
                TriPos tp;
                __builtin_memcpy(&tp, &v[3 * j], sizeof(TriPos));
                float t, u, vv;
                if (tri_test(tp, o, tmin, fminf_(tmax, h.t), &t, &u, &vv) && t < h.t) {
                    h.t = t; h.slot = (int32_t)(la + j); h.inst = cur_inst; h.u = u; h.v = vv;
                }
            }
        }
        if (lb > (uint32_t)TRI_BATCH) { la += TRI_BATCH; lb -= TRI_BATCH; return true; }
        lb = 0u;
        return next();
    }
};

// Any-hit traversal state of one lane (entries and blocks as above; no
// closest-distance bound, nearer child first).
struct AnyTrav {
    Ray w, o;
    float tmin, tmax;
    bool hit;
    uint32_t cur;
    uint32_t la, lb;
    LaneStack st;

    KD bool init(const SceneView& S, const Ray& ray, float t0, float t1, const StackShape& sh, uint32_t* spill) {
        w = ray; o = ray; tmin = t0; tmax = t1; hit = false;
        la = 0u; lb = 0u;
        st.reset(sh, spill);
        if (S.tlas_root < 0) return false;
        float4 rlo, rhi;
        load_node(S.tlas_nodes, 0, &rlo, &rhi);
        float te;
        if (!slab(rlo, rhi, w, tmin, tmax, &te)) return false;
        cur = tlas_entry(bits_u(rlo.w), bits_u(rhi.w));
        return true;
    }
    KD bool next() {
        if (st.sp == 0 || st.overflow) return false;
        cur = st.pop();
        return true;
    }
    KD bool children(bool tl, const float4& l1, const float4& h1, const float4& l2, const float4& h2, uint32_t i1,
                     uint32_t i2, int* near_k) {
        Ray r;
        r.o = tl ? w.o : o.o; r.d = tl ? w.d : o.d; r.inv = tl ? w.inv : o.inv;
        float c1, c2;
        const bool k1 = slab(l1, h1, r, tmin, tmax, &c1);
        const bool k2 = slab(l2, h2, r, tmin, tmax, &c2);
        if (k1 && k2) {   // nearer first: an occluder near the origin ends the query sooner
            const bool far1 = c1 > c2;
            const uint32_t fa = bits_u(far1 ? l1.w : l2.w), fb = bits_u(far1 ? h1.w : h2.w);
            st.push(tl ? tlas_entry(fa, fb) : blas_entry(far1 ? i1 : i2, fa, fb));
            *near_k = far1 ? 1 : 0;
            return true;
        }
        if (k1) { *near_k = 0; return true; }
        if (k2) { *near_k = 1; return true; }
        return false;
    }
    template <bool COUNT>
    KD bool step(const SceneView& S, TraceCounts* cnt) {
        const uint32_t e = cur;
        const bool new_leaf = (lb == 0u) && (e & LEAF_BIT) && !(e & TLAS_BIT);
        const float4* base;
        int n;
        const int kind = step_kind(S, e, la, lb, &base, &n);
        float4 v[STEP_LOADS];
        step_load(base, n, v);
        if (kind == STEP_INTERIOR) {
            const bool tl = (e & TLAS_BIT) != 0;
            const uint32_t q = e & ENTRY_PAYLOAD;
            int nk;
            bool any;
            if (COUNT) { if (tl) cnt->tlas++; else cnt->interior++; }
            any = children(tl, v[0], v[1], v[2], v[3], 2 * q, 2 * q + 1, &nk);
            if (!any) return next();
            if (st.overflow) return false;
            const float4 nl = nk ? v[2] : v[0], nh = nk ? v[3] : v[1];
            const uint32_t na = bits_u(nl.w), nb = bits_u(nh.w);
            const bool n_interior = tl ? (na != 0u) : (nb == 0u);
            if (!n_interior) {
                cur = tl ? tlas_entry(na, nb) : blas_entry(2 * q + (uint32_t)nk, na, nb);
                return true;
            }
            const float4 al = nk ? v[8] : v[4], ah = nk ? v[9] : v[5];
            const float4 bl = nk ? v[10] : v[6], bh = nk ? v[11] : v[7];
            const uint32_t p = na >> 1;
            if (COUNT) { if (tl) cnt->tlas++; else cnt->interior++; }
            any = children(tl, al, ah, bl, bh, 2 * p, 2 * p + 1, &nk);
            if (!any) return next();
            if (st.overflow) return false;
            const float4 ml = nk ? bl : al, mh = nk ? bh : ah;
            cur = tl ? tlas_entry(bits_u(ml.w), bits_u(mh.w)) : blas_entry(2 * p + (uint32_t)nk, bits_u(ml.w), bits_u(mh.w));
            return true;
        }
        if (kind == STEP_INSTANCE) {
            if (COUNT) { cnt->tlas++; cnt->entries++; }
            InstTrav it;
            __builtin_memcpy(&it, v, sizeof(InstTrav));
            o = object_ray(it, w);
            if (it.root_entry != 0xFFFFFFFFu) { cur = it.root_entry; return true; }
            return next();
        }
        if (kind == STEP_HEADER) {
            la = bits_u(v[0].w);
            lb = bits_u(v[1].w);
            if (COUNT) cnt->leaf++;
            return true;
        }
        if (COUNT && new_leaf) cnt->leaf++;
#pragma unroll
        for (int j = 0; j < TRI_BATCH; ++j) {
            if ((uint32_t)j < lb) {
                if (COUNT) cnt->tri++;
                TriPos tp;
                __builtin_memcpy(&tp, &v[3 * j], sizeof(TriPos));
                float t, u, vv;
                if (tri_test(tp, o, tmin, tmax, &t, &u, &vv)) { hit = true; lb = 0u; return false; }
            }
        }
        if (lb > (uint32_t)TRI_BATCH) { la += TRI_BATCH; lb -= TRI_BATCH; return true; }
        lb = 0u;
        return next();
    }
};

// Whole-query forms (megakernel, trace entry points, the wavefront's tail).
template <bool COUNT>
KD Hit trace_closest(const SceneView& S, const Ray& w, float tmin, float tmax, const StackShape& sh,
                     bool* overflow, TraceCounts* cnt) {
    uint32_t spill[SPILL_STACK];
    ClosestTrav tr;
    if (tr.init(S, w, tmin, tmax, sh, spill)) {
        while (tr.template step<COUNT>(S, cnt)) {}
    }
    *overflow = tr.st.overflow;
    return tr.h;
}

template <bool COUNT>
KD bool trace_any_in(const SceneView& S, const Ray& w, float tmin, float tmax, const StackShape& sh, uint32_t* spill,
                     bool* overflow, TraceCounts* cnt) {
    AnyTrav tr;
    if (tr.init(S, w, tmin, tmax, sh, spill)) {
        while (tr.template step<COUNT>(S, cnt)) {}
    }
    *overflow = tr.st.overflow;
    return tr.hit;
}

template <bool COUNT>
KD bool trace_any(const SceneView& S, const Ray& w, float tmin, float tmax, const StackShape& sh, bool* overflow,
                  TraceCounts* cnt) {
    uint32_t spill[SPILL_STACK];
    return trace_any_in<COUNT>(S, w, tmin, tmax, sh, spill, overflow, cnt);
}

// XCD-aware workgroup swizzle: the dispatcher hands consecutive workgroup ids
// to the 8 XCDs round-robin; remap so each XCD owns a contiguous run of tiles
// (neighbouring tiles share BVH nodes in that XCD's L2).
KD uint32_t xcd_swizzle(uint32_t bid, uint32_t nblocks) {
    const uint32_t X = 8;
    uint32_t per = nblocks / X;
    if (per == 0 || bid >= per * X) return bid;
    return (bid % X) * per + (bid / X);
}

}  // namespace kpt
