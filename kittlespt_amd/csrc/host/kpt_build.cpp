// kpt_build.cpp — BLAS (binned SAH) and TLAS (agglomerative) builders that
// reproduce the reference's trees node for node, written as explicit-stack
// loops over flat arrays.
#include "kpt_build.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <utility>

namespace kpt {
namespace {

struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow_pt(const float* p) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::fmin(lo[k], p[k]); hi[k] = std::fmax(hi[k], p[k]); }
    }
    // Bounds3f::grow(Bounds3f): pmax first, then pmin (bounds.cuh:36-40)
    void grow_box(const Box& b) { grow_pt(b.hi); grow_pt(b.lo); }
    // Bounds3f::surfaceArea (bounds.cuh:26-30)
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

inline float axis_of(const float* v, int axis) { return axis == 0 ? v[0] : axis == 1 ? v[1] : v[2]; }

struct BlasWork {
    const float* pos;
    std::vector<float> cen;        // 3 per triangle of the mesh (local index)
    int32_t first;
    std::vector<int32_t> idx;      // slot -> local triangle index
    struct Tmp { int32_t first; uint32_t count; Box box; };
    std::vector<Tmp> tmp;

    const float* vtx(int32_t local, int v) const { return pos + 9 * (size_t)(first + local) + 3 * v; }

    Box box_of(int32_t start, uint32_t count) const {
        Box b;
        for (uint32_t i = 0; i < count; ++i) {
            int32_t t = idx[start + i];
            b.grow_pt(vtx(t, 0)); b.grow_pt(vtx(t, 1)); b.grow_pt(vtx(t, 2));
        }
        return b;
    }

    // findBestSplitPlane (as_builder.cuh:110-179)
    void best_split(int32_t start, uint32_t count, int* axis, float* split) const {
        float best = INFINITY;
        for (int a = 0; a < 3; ++a) {
            float bmin = INFINITY, bmax = -INFINITY;
            for (uint32_t i = 0; i < count; ++i) {
                float c = cen[3 * (size_t)idx[start + i] + a];
                bmin = std::fmin(bmin, c);
                bmax = std::fmax(bmax, c);
            }
            if (bmin == bmax) continue;
            Box bins[8];
            uint32_t bc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            float scale = 8.0f / (bmax - bmin);
            for (uint32_t i = 0; i < count; ++i) {
                int32_t t = idx[start + i];
                // min(BINS - 1, (int)((c - min) * scale)) for every finite
                // centroid; a non-finite one (NaN, or inf * 0 when another
                // centroid is infinite) goes to bin 0 instead of an undefined
                // float-to-int conversion
                const float f = (cen[3 * (size_t)t + a] - bmin) * scale;
                int b = f >= 0.0f ? (int)std::fmin(7.0f, f) : 0;
                bc[b]++;
                bins[b].grow_pt(vtx(t, 0)); bins[b].grow_pt(vtx(t, 1)); bins[b].grow_pt(vtx(t, 2));
            }
            float la[7], ra[7];
            int32_t lc[7], rc[7];
            Box lb, rb;
            int32_t ls = 0, rs = 0;
            for (int i = 0; i < 7; ++i) {
                ls += bc[i]; lc[i] = ls; lb.grow_box(bins[i]); la[i] = lb.area();
                rs += bc[7 - i]; rc[6 - i] = rs; rb.grow_box(bins[7 - i]); ra[6 - i] = rb.area();
            }
            float step = (bmax - bmin) / 8.0f;
            for (int i = 0; i < 7; ++i) {
                float cost = lc[i] * la[i] + rc[i] * ra[i];
                if (cost < best) { *axis = a; *split = bmin + step * (i + 1); best = cost; }
            }
        }
    }
};

void to_node(Node& n, const Box& b, uint32_t a, uint32_t cnt) {
    for (int k = 0; k < 3; ++k) { n.lo[k] = b.lo[k]; n.hi[k] = b.hi[k]; }
    n.a = a;
    n.b = cnt;
}

}  // namespace

BlasTree build_blas(const float* pos, int32_t first, int32_t count) {
    BlasTree out;
    if (count <= 0) return out;
    BlasWork w;
    w.pos = pos;
    w.first = first;
    w.cen.resize(3 * (size_t)count);
    for (int32_t t = 0; t < count; ++t) {
        const float* p0 = w.vtx(t, 0); const float* p1 = w.vtx(t, 1); const float* p2 = w.vtx(t, 2);
        for (int k = 0; k < 3; ++k) w.cen[3 * (size_t)t + k] = (p0[k] + p1[k] + p2[k]) * 0.3333f;
    }
    w.idx.resize(count);
    for (int32_t i = 0; i < count; ++i) w.idx[i] = i;
    const size_t cap = 2 * (size_t)count - 1;
    w.tmp.assign(cap, BlasWork::Tmp{-1, 0, Box()});
    w.tmp[0] = BlasWork::Tmp{0, (uint32_t)count, w.box_of(0, count)};
    std::vector<int32_t> child(cap, -1);   // interior: left child local index
    int32_t next = 1;
    // pre-order processing == the reference's recursion (left subtree first)
    std::vector<std::pair<int32_t, int32_t>> todo;   // (node, depth)
    todo.push_back({0, 0});
    while (!todo.empty()) {
        auto [nid, depth] = todo.back();
        todo.pop_back();
        out.max_depth = std::max(out.max_depth, depth);
        BlasWork::Tmp& node = w.tmp[nid];
        if (node.count <= 8) continue;
        int axis = -1;
        float split = 0.0f;
        w.best_split(node.first, node.count, &axis, &split);
        int32_t i = node.first, j = i + (int32_t)node.count - 1;
        while (i <= j) {
            if (axis_of(&w.cen[3 * (size_t)w.idx[i]], axis) < split) i++;
            else std::swap(w.idx[i], w.idx[j--]);
        }
        uint32_t lcount = (uint32_t)(i - node.first);
        if (lcount == 0 || lcount == node.count) continue;
        int32_t l = next++, r = next++;
        w.tmp[l] = BlasWork::Tmp{node.first, lcount, w.box_of(node.first, lcount)};
        w.tmp[r] = BlasWork::Tmp{i, node.count - lcount, w.box_of(i, node.count - lcount)};
        child[nid] = l;
        w.tmp[nid].count = 0;
        todo.push_back({r, depth + 1});
        todo.push_back({l, depth + 1});
    }
    out.used = next;
    out.nodes.resize(cap);
    std::memset(out.nodes.data(), 0, cap * sizeof(Node));
    for (int32_t n = 0; n < next; ++n) {
        const BlasWork::Tmp& t = w.tmp[n];
        if (t.count == 0) to_node(out.nodes[n], t.box, (uint32_t)child[n], 0u);
        else to_node(out.nodes[n], t.box, (uint32_t)t.first, t.count);
    }
    out.order.resize(count);
    for (int32_t s = 0; s < count; ++s) out.order[s] = first + w.idx[s];
    const Box& rb = w.tmp[0].box;
    for (int k = 0; k < 3; ++k) { out.bounds.lo[k] = rb.lo[k]; out.bounds.hi[k] = rb.hi[k]; }
    return out;
}

Aabb transform_bounds(const Aabb& b, const float M[16]) {
    Box o;
    for (int i = 0; i < 8; ++i) {
        float v[4] = {(i & 1) ? b.hi[0] : b.lo[0], (i & 2) ? b.hi[1] : b.lo[1], (i & 4) ? b.hi[2] : b.lo[2], 1.0f};
        float p[3];
        for (int r = 0; r < 3; ++r) p[r] = M[0 * 4 + r] * v[0] + M[1 * 4 + r] * v[1] + M[2 * 4 + r] * v[2] + M[3 * 4 + r] * v[3];
        o.grow_pt(p);
    }
    Aabb a;
    for (int k = 0; k < 3; ++k) { a.lo[k] = o.lo[k]; a.hi[k] = o.hi[k]; }
    return a;
}

TlasTree build_tlas(const std::vector<Aabb>& inst) {
    TlasTree out;
    const int32_t N = (int32_t)inst.size();
    if (N == 0) return out;
    std::vector<Box> box(2 * (size_t)N);
    std::vector<uint32_t> lr(2 * (size_t)N, 0u);
    std::vector<int32_t> blas(2 * (size_t)N, -1);
    std::vector<int32_t> ids(N);
    int32_t ptr = 1;
    for (int32_t i = 0; i < N; ++i) {
        ids[i] = ptr;
        for (int k = 0; k < 3; ++k) { box[ptr].lo[k] = inst[i].lo[k]; box[ptr].hi[k] = inst[i].hi[k]; }
        blas[ptr] = i;
        lr[ptr++] = 0u;
    }
    // Active boxes in list order (structure of arrays, mirroring ids[]) so the
    // best-match scan streams through memory and vectorises; the arithmetic per
    // candidate is the reference's, and the scan keeps its first minimum.
    std::vector<float> slo[3], shi[3], sa_buf(N);
    for (int k = 0; k < 3; ++k) {
        slo[k].resize(N);
        shi[k].resize(N);
        for (int32_t i = 0; i < N; ++i) { slo[k][i] = box[ids[i]].lo[k]; shi[k][i] = box[ids[i]].hi[k]; }
    }
    auto fmaxf_ = [](float a, float b) { return a > b ? a : b; };   // finite inputs: fmax up to the sign of 0
    auto fminf_ = [](float a, float b) { return a < b ? a : b; };
    // findBestMatch: smallest half area of the merged box, first minimum wins
    auto best_match = [&](int32_t n, int32_t A) {
        const float ax0 = slo[0][A], ay0 = slo[1][A], az0 = slo[2][A];
        const float ax1 = shi[0][A], ay1 = shi[1][A], az1 = shi[2][A];
        const float *l0 = slo[0].data(), *l1 = slo[1].data(), *l2 = slo[2].data();
        const float *h0 = shi[0].data(), *h1 = shi[1].data(), *h2 = shi[2].data();
        float* sa = sa_buf.data();
        for (int32_t B = 0; B < n; ++B) {
            const float ex = fmaxf_(ax1, h0[B]) - fminf_(ax0, l0[B]);
            const float ey = fmaxf_(ay1, h1[B]) - fminf_(ay0, l1[B]);
            const float ez = fmaxf_(az1, h2[B]) - fminf_(az0, l2[B]);
            sa[B] = ex * ey + ey * ez + ez * ex;
        }
        float smallest = INFINITY;
        int32_t best = -1;
        for (int32_t B = 0; B < n; ++B) {
            if (B == A) continue;
            if (sa[B] < smallest) { smallest = sa[B]; best = B; }
        }
        // No merged area below infinity (non-finite or overflowing boxes):
        // the reference returns -1 and then indexes with it; here the first
        // other entry is taken, as if every such area were one equal value.
        if (best < 0 && n > 1) best = A == 0 ? 1 : 0;
        return best;
    };
    int32_t left = N;
    int32_t A = 0, B = best_match(left, A);
    // The walk merges N - 1 times and moves on a few times per merge; past
    // the cap (never reached for finite boxes) A and B are merged as they
    // stand, so a degenerate input still ends with a tree.
    const int64_t max_steps = 64 * (int64_t)N + 1024;
    for (int64_t steps = 0; left > 1; ++steps) {
        int32_t C = steps > max_steps ? A : best_match(left, B);
        if (A == C) {
            int32_t na = ids[A], nb = ids[B];
            Box m;
            for (int k = 0; k < 3; ++k) { m.lo[k] = std::fmin(box[na].lo[k], box[nb].lo[k]); m.hi[k] = std::fmax(box[na].hi[k], box[nb].hi[k]); }
            box[ptr] = m;
            lr[ptr] = (uint32_t)na + ((uint32_t)nb << 16);
            ids[A] = ptr++;
            ids[B] = ids[left - 1];
            for (int k = 0; k < 3; ++k) {
                slo[k][A] = m.lo[k]; shi[k][A] = m.hi[k];
                slo[k][B] = slo[k][left - 1]; shi[k][B] = shi[k][left - 1];
            }
            B = best_match(--left, A);
        } else {
            A = B;
            B = C;
        }
    }
    box[0] = box[ids[A]];
    lr[0] = lr[ids[A]];
    blas[0] = blas[ids[A]];
    out.root = 0;
    out.nodes.resize(ptr);
    for (int32_t n = 0; n < ptr; ++n) to_node(out.nodes[n], box[n], lr[n], (uint32_t)blas[n]);
    // depth statistic
    std::vector<std::pair<uint32_t, int>> st{{0u, 0}};
    while (!st.empty()) {
        auto [n, d] = st.back();
        st.pop_back();
        out.max_depth = std::max(out.max_depth, d);
        if (lr[n] != 0u) { st.push_back({lr[n] & 0xFFFFu, d + 1}); st.push_back({(lr[n] >> 16) & 0xFFFFu, d + 1}); }
    }
    return out;
}

}  // namespace kpt
