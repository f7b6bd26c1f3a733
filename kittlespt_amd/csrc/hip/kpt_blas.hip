// kpt_blas.hip — BLAS builds on the device that produce the host builder's
// trees node for node (BLASBuilder, as_builder.cuh:110-261: centroid
// (p0+p1+p2)*0.3333f, 8 bins per axis over the centroid extent, SAH sweep,
// leaves of <= 8 triangles, the in-place two-pointer partition, abort on an
// empty side, children allocated as pairs in pre-order).
//
// All meshes are built together, one level of the tree per round:
//   bounds  centroid extent of every node being split (orderable-key atomics)
//   bins    8 bins x 3 axes of counts and boxes (orderable-key atomics)
//   split   the serial SAH sweep, one thread per node, the host's operation
//           order exactly
//   flag    1 where a triangle goes right; one exclusive scan over all slots
//   scatter the permutation of the reference's swap partition in closed form:
//           with L lefts in a node of n slots and A = L + (slot L goes right),
//           slots r < A that go left stay; the k-th right among them moves to
//           n-1 (k = 0) or one below the (k-1)-th left of the back part
//           [A, n) counted from the end; the k-th such back left takes the
//           k-th front right's slot; back rights move down by one
//   box     child boxes with the host's fminf/fmaxf tie rule (the later
//           vertex wins between equal values, e.g. +0 and -0): 64-bit keys
//           (value key << 32 | slot position) so that atomics pick the value
//           and the position of the winner, decoded from the vertex itself
// Nodes get breadth-first ids while they are built; a final pass counts the
// split nodes of every subtree and renumbers them in the reference's
// pre-order pair allocation: the children of the split node with pre-order
// rank q are 1+2q and 2+2q.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <cstdint>
#include <cstring>
#include <vector>

#include "kpt_launch.h"

namespace kpt {
namespace {

constexpr int BT = 256;
constexpr int LEAF_MAX = 8;
constexpr int NBINS = 8;
constexpr int BIN_WORDS = 3 * NBINS * 7;   // counts, then lo keys, then hi keys (axis*8+bin)*3+comp
constexpr uint64_t LOW = 0xFFFFFFFFull;

// Orderable key of a float: unsigned order = value order, +0 and -0 share
// the key of +0 (they compare equal, the tie rule decides between them).
__device__ __forceinline__ uint32_t okey(float v) {
    uint32_t u = __float_as_uint(v);
    if (v == 0.0f) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float okey_value(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}
__device__ __forceinline__ float comp3(const float* v, int axis) { return axis == 0 ? v[0] : axis == 1 ? v[1] : v[2]; }

struct BB {
    int32_t T;
    const float* pos;        // 9 floats per triangle, global ids
    float* cen;              // 3 per triangle
    int32_t* idx;            // slot -> triangle
    int32_t* idx2;           // scatter target
    int32_t* slot_act;       // slot -> index of its node in this level's split list, -1 if settled
    int32_t* flag;           // T+1, 1 = goes right
    int32_t* scan;           // T+1, exclusive sums of flag
    int32_t* fr;             // per node range: offsets of the front rights
    int32_t* bl;             // per node range: offsets of the back lefts
    // nodes, breadth-first ids (mesh m's root is node m)
    int32_t* nfirst; int32_t* ncount; int32_t* nchild; int32_t* nmesh; int32_t* ndepth; int32_t* nnext;
    unsigned long long* nkey;   // 6 per node: lo keys, hi keys
    float* nbox;                // 6 per node: lo, hi
    // nodes being split at this level
    const int32_t* act;
    int32_t* act_next;
    uint32_t* acb;           // 6 per entry: centroid min keys, max keys
    uint32_t* abin;          // BIN_WORDS per entry
    int32_t* aaxis; float* asplit; int32_t* aleft;
    int32_t* counters;       // [0] nodes allocated, [1] next level's split-list length, [2] error, [3] tasks
    int32_t node_cap, act_cap, task_cap;
    int2* task;              // this level's (entry, chunk) tasks
    int32_t* mesh_depth;
    const int32_t* mesh_first; int32_t M;
};

__global__ __launch_bounds__(BT) void k_bb_init(BB b, const int32_t* root_act) {
    const int32_t s = blockIdx.x * BT + threadIdx.x;
    if (s == b.T) b.flag[s] = 0;
    if (s >= b.T) return;
    const float* p = b.pos + 9 * (size_t)s;
    for (int k = 0; k < 3; ++k) b.cen[3 * (size_t)s + k] = (p[k] + p[3 + k] + p[6 + k]) * 0.3333f;
    b.idx[s] = s;
    int lo = 0, hi = b.M;                       // last mesh with first <= s
    while (hi - lo > 1) { int mid = (lo + hi) >> 1; if (b.mesh_first[mid] <= s) lo = mid; else hi = mid; }
    b.slot_act[s] = root_act[lo];
}

// One thread per bin word of the level's nodes: nact * BIN_WORDS passes 2^31
// once a level splits more than ~12.8 M nodes (meshes of ~115 M triangles),
// so the index is 64-bit.
__global__ __launch_bounds__(BT) void k_bb_level_init(BB b, int32_t nact) {
    const int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i == 0) { b.counters[1] = 0; b.counters[3] = 0; }
    if (i < (int64_t)nact * 6) b.acb[i] = (i % 6) < 3 ? okey(INFINITY) : okey(-INFINITY);
    if (i < (int64_t)nact * BIN_WORDS) {
        const int w = (int)(i % BIN_WORDS);
        b.abin[i] = w < 24 ? 0u : w < 96 ? okey(FLT_MAX) : okey(-FLT_MAX);
    }
}

// Work of the reductions comes in tasks of up to TASK consecutive slots of
// one node, one wave each: per-lane partials, a wave reduction, then one
// atomic per value and task (a few hundred per address at the top levels
// instead of one per lane).
constexpr int TW = 64;
constexpr int TASK = 512;

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) { const uint32_t w = (uint32_t)__shfl_xor((int)v, o); v = w < v ? w : v; }
    return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) { const uint32_t w = (uint32_t)__shfl_xor((int)v, o); v = w > v ? w : v; }
    return v;
}
__device__ __forceinline__ unsigned long long shfl_xor64(unsigned long long v, int o) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o), hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o);
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned long long wave_min64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) { const unsigned long long w = shfl_xor64(v, o); v = w < v ? w : v; }
    return v;
}
__device__ __forceinline__ unsigned long long wave_max64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) { const unsigned long long w = shfl_xor64(v, o); v = w > v ? w : v; }
    return v;
}

// Tasks of this level's split list: (entry, chunk) pairs.
__global__ __launch_bounds__(BT) void k_bb_tasks(BB b, int32_t nact) {
    const int32_t a = blockIdx.x * BT + threadIdx.x;
    if (a >= nact) return;
    const int32_t count = b.ncount[b.act[a]], nt = (count + TASK - 1) / TASK;
    const int32_t off = atomicAdd(&b.counters[3], nt);
    if (off + nt > b.task_cap) { b.counters[2] = 1; return; }
    for (int32_t j = 0; j < nt; ++j) b.task[off + j] = make_int2(a, j);
}

// Node boxes.  mode 0: tasks (mesh, chunk) over the roots; mode 1: tasks
// (entry, chunk) over this level's splits, boxes of their two children, and
// each slot's entry in the next level's split list.
__global__ __launch_bounds__(TW) void k_bb_box(BB b, int mode, const int2* tasks, int32_t ntasks) {
    const int32_t t = blockIdx.x;
    if (t >= (mode == 0 ? ntasks : b.counters[3])) return;
    const int2 tk = tasks[t];
    const int32_t n = mode == 0 ? tk.x : b.act[tk.x];
    const int32_t first = b.nfirst[n], count = b.ncount[n];
    const int32_t c = mode == 0 ? n : b.nchild[n];
    const int32_t L = mode == 0 ? count : b.aleft[tk.x];
    const int32_t beg = tk.y * TASK, end = min(count, beg + TASK);
    unsigned long long lo[2][3], hi[2][3];
    for (int h = 0; h < 2; ++h) for (int k = 0; k < 3; ++k) { lo[h][k] = ~0ull; hi[h][k] = 0ull; }
    for (int32_t r = beg + (int32_t)threadIdx.x; r < end; r += TW) {
        const int32_t s = first + r;
        if (mode == 1) {
            if (c < 0) { b.slot_act[s] = -1; continue; }
            b.slot_act[s] = b.nnext[c + (r >= L ? 1 : 0)];
        }
        const int h = r >= L ? 1 : 0;
        const float* p = b.pos + 9 * (size_t)b.idx[s];
        for (int v = 0; v < 3; ++v) {
            const uint64_t q = (uint64_t)(3u * (uint32_t)s + (uint32_t)v + 1u);
            for (int k = 0; k < 3; ++k) {
                const float x = p[3 * v + k];
                if (x != x) continue;
                const uint64_t kk = (uint64_t)okey(x) << 32;
                const unsigned long long kl = kk | (LOW - q), kh = kk | q;
                lo[h][k] = kl < lo[h][k] ? kl : lo[h][k];
                hi[h][k] = kh > hi[h][k] ? kh : hi[h][k];
            }
        }
    }
    if (c < 0) return;
    for (int h = 0; h < (mode == 0 ? 1 : 2); ++h) {
        for (int k = 0; k < 3; ++k) {
            const unsigned long long l = wave_min64(lo[h][k]), u = wave_max64(hi[h][k]);
            if (threadIdx.x == 0) {
                const int32_t tn = c + h;
                if (l != ~0ull) atomicMin(&b.nkey[6 * (size_t)tn + k], l);
                if (u != 0ull) atomicMax(&b.nkey[6 * (size_t)tn + 3 + k], u);
            }
        }
    }
}

// Box floats of nodes [lo, counters[0]) (or [lo, hi) when hi >= 0) from
// their keys: FLT_MAX / -FLT_MAX when no vertex beat the initial value,
// else the winning vertex's own value.
__global__ __launch_bounds__(BT) void k_bb_decode(BB b, int32_t lo, int32_t hi) {
    const int32_t n = lo + blockIdx.x * BT + threadIdx.x;
    const int32_t end = hi >= 0 ? hi : b.counters[0];
    if (n >= end) return;
    for (int k = 0; k < 6; ++k) {
        const unsigned long long key = b.nkey[6 * (size_t)n + k];
        const uint64_t low = key & LOW;
        float v;
        if (k < 3 && low == LOW) v = FLT_MAX;
        else if (k >= 3 && low == 0) v = -FLT_MAX;
        else {
            const uint64_t q = (k < 3 ? LOW - low : low) - 1;
            const uint32_t s = (uint32_t)(q / 3), vi = (uint32_t)(q % 3);
            if (s >= (uint32_t)b.T) { b.counters[2] = 1; v = 0.0f; }
            else v = b.pos[9 * (size_t)b.idx[s] + 3 * vi + (k % 3)];
        }
        b.nbox[6 * (size_t)n + k] = v;
    }
}

__global__ __launch_bounds__(TW) void k_bb_bounds(BB b) {
    const int32_t t = blockIdx.x;
    if (t >= b.counters[3]) return;
    const int2 tk = b.task[t];
    const int32_t a = tk.x, n = b.act[a], first = b.nfirst[n], count = b.ncount[n];
    const int32_t beg = tk.y * TASK, end = min(count, beg + TASK);
    uint32_t mn[3] = {~0u, ~0u, ~0u}, mx[3] = {0u, 0u, 0u};
    for (int32_t r = beg + (int32_t)threadIdx.x; r < end; r += TW) {
        const float* cp = b.cen + 3 * (size_t)b.idx[first + r];
        for (int k = 0; k < 3; ++k) { const uint32_t c = okey(cp[k]); mn[k] = c < mn[k] ? c : mn[k]; mx[k] = c > mx[k] ? c : mx[k]; }
    }
    for (int k = 0; k < 3; ++k) {
        const uint32_t l = wave_min(mn[k]), u = wave_max(mx[k]);
        if (threadIdx.x == 0) { atomicMin(&b.acb[6 * (size_t)a + k], l); atomicMax(&b.acb[6 * (size_t)a + 3 + k], u); }
    }
}

// Bin boxes only feed surface areas, where the sign of a zero bound cannot
// change a product's comparison, so value-only keys suffice here.
__global__ __launch_bounds__(TW) void k_bb_bins(BB b) {
    __shared__ uint32_t sb[BIN_WORDS];
    const int32_t t = blockIdx.x;
    if (t >= b.counters[3]) return;
    const int2 tk = b.task[t];
    const int32_t a = tk.x, n = b.act[a], first = b.nfirst[n], count = b.ncount[n];
    const int32_t beg = tk.y * TASK, end = min(count, beg + TASK);
    for (int w = threadIdx.x; w < BIN_WORDS; w += TW) sb[w] = w < 24 ? 0u : w < 96 ? okey(FLT_MAX) : okey(-FLT_MAX);
    float bmin[3], scale[3];
    bool live[3];
    for (int ax = 0; ax < 3; ++ax) {
        bmin[ax] = okey_value(b.acb[6 * (size_t)a + ax]);
        const float bmax = okey_value(b.acb[6 * (size_t)a + 3 + ax]);
        live[ax] = bmin[ax] != bmax;
        scale[ax] = live[ax] ? 8.0f / (bmax - bmin[ax]) : 0.0f;
    }
    __syncthreads();
    for (int32_t r = beg + (int32_t)threadIdx.x; r < end; r += TW) {
        const int32_t tri = b.idx[first + r];
        const float* p = b.pos + 9 * (size_t)tri;
        uint32_t tlo[3], thi[3];
        for (int k = 0; k < 3; ++k) {
            float l = p[k], h = p[k];
            for (int v = 1; v < 3; ++v) { const float x = p[3 * v + k]; l = x < l ? x : l; h = x > h ? x : h; }
            tlo[k] = okey(l); thi[k] = okey(h);
        }
        const float* cp = b.cen + 3 * (size_t)tri;
        for (int ax = 0; ax < 3; ++ax) {
            if (!live[ax]) continue;
            const int32_t raw = (int32_t)((cp[ax] - bmin[ax]) * scale[ax]);
            const int e = ax * NBINS + (raw < NBINS - 1 ? raw : NBINS - 1);
            atomicAdd(&sb[e], 1u);
            for (int k = 0; k < 3; ++k) { atomicMin(&sb[24 + 3 * e + k], tlo[k]); atomicMax(&sb[96 + 3 * e + k], thi[k]); }
        }
    }
    __syncthreads();
    uint32_t* g = b.abin + (size_t)a * BIN_WORDS;
    for (int w = threadIdx.x; w < BIN_WORDS; w += TW) {
        const int e = w < 24 ? w : w < 96 ? (w - 24) / 3 : (w - 96) / 3;
        if (sb[e] == 0u) continue;                    // an empty bin keeps its initial box
        if (w < 24) atomicAdd(&g[w], sb[w]);
        else if (w < 96) atomicMin(&g[w], sb[w]);
        else atomicMax(&g[w], sb[w]);
    }
}

// Host Box semantics for the sweep (kpt_build.cpp / bounds.cuh:20-40).
struct SBox {
    float lo[3], hi[3];
    __device__ void init() { for (int k = 0; k < 3; ++k) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; } }
    __device__ void grow_pt(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = (lo[k] < p[k] || p[k] != p[k]) ? lo[k] : p[k];
            hi[k] = (hi[k] > p[k] || p[k] != p[k]) ? hi[k] : p[k];
        }
    }
    __device__ void grow_box(const SBox& o) { grow_pt(o.hi); grow_pt(o.lo); }
    __device__ float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

// findBestSplitPlane (as_builder.cuh:110-179), one thread per node.
__global__ __launch_bounds__(BT) void k_bb_split(BB b, int32_t nact) {
    const int32_t a = blockIdx.x * BT + threadIdx.x;
    if (a >= nact) return;
    const uint32_t* g = b.abin + (size_t)a * BIN_WORDS;
    float best = INFINITY, split = 0.0f;
    int axis = -1;
    for (int ax = 0; ax < 3; ++ax) {
        const float bmin = okey_value(b.acb[6 * (size_t)a + ax]), bmax = okey_value(b.acb[6 * (size_t)a + 3 + ax]);
        if (bmin == bmax) continue;
        SBox bins[NBINS];
        int32_t bc[NBINS];
        for (int i = 0; i < NBINS; ++i) {
            const int e = ax * NBINS + i;
            bc[i] = (int32_t)g[e];
            for (int k = 0; k < 3; ++k) { bins[i].lo[k] = okey_value(g[24 + 3 * e + k]); bins[i].hi[k] = okey_value(g[96 + 3 * e + k]); }
        }
        float la[7], ra[7];
        int32_t lc[7], rc[7];
        SBox lb, rb;
        lb.init(); rb.init();
        int32_t ls = 0, rs = 0;
        for (int i = 0; i < 7; ++i) {
            ls += bc[i]; lc[i] = ls; lb.grow_box(bins[i]); la[i] = lb.area();
            rs += bc[7 - i]; rc[6 - i] = rs; rb.grow_box(bins[7 - i]); ra[6 - i] = rb.area();
        }
        const float step = (bmax - bmin) / 8.0f;
        for (int i = 0; i < 7; ++i) {
            const float cost = (float)lc[i] * la[i] + (float)rc[i] * ra[i];
            if (cost < best) { axis = ax; split = bmin + step * (float)(i + 1); best = cost; }
        }
    }
    b.aaxis[a] = axis;
    b.asplit[a] = split;
}

__global__ __launch_bounds__(BT) void k_bb_flag(BB b) {
    const int32_t s = blockIdx.x * BT + threadIdx.x;
    if (s >= b.T) return;
    const int32_t a = b.slot_act[s];
    int32_t f = 0;
    if (a >= 0) f = comp3(b.cen + 3 * (size_t)b.idx[s], b.aaxis[a]) < b.asplit[a] ? 0 : 1;
    b.flag[s] = f;
}

// Left counts; children for splits with both sides non-empty.
__global__ __launch_bounds__(BT) void k_bb_left(BB b, int32_t nact) {
    const int32_t a = blockIdx.x * BT + threadIdx.x;
    if (a >= nact) return;
    const int32_t n = b.act[a], first = b.nfirst[n], count = b.ncount[n];
    const int32_t L = count - (b.scan[first + count] - b.scan[first]);
    b.aleft[a] = L;
    if (L == 0 || L == count) { b.nchild[n] = -1; return; }
    const int32_t c = atomicAdd(&b.counters[0], 2);
    if (c + 2 > b.node_cap) { b.counters[2] = 1; b.nchild[n] = -1; return; }
    b.nchild[n] = c;
    const int32_t d = b.ndepth[n] + 1, m = b.nmesh[n];
    const int32_t f2[2] = {first, first + L}, n2[2] = {L, count - L};
    for (int h = 0; h < 2; ++h) {
        const int32_t ch = c + h;
        b.nfirst[ch] = f2[h]; b.ncount[ch] = n2[h]; b.nchild[ch] = -1; b.nmesh[ch] = m; b.ndepth[ch] = d;
        for (int k = 0; k < 3; ++k) { b.nkey[6 * (size_t)ch + k] = ((uint64_t)okey(FLT_MAX) << 32) | LOW; b.nkey[6 * (size_t)ch + 3 + k] = (uint64_t)okey(-FLT_MAX) << 32; }
        int32_t nx = -1;
        if (n2[h] > LEAF_MAX) {
            nx = atomicAdd(&b.counters[1], 1);
            if (nx < b.act_cap) b.act_next[nx] = ch;
            else { b.counters[2] = 1; nx = -1; }
        }
        b.nnext[ch] = nx;
    }
    atomicMax(&b.mesh_depth[m], d);
}

__device__ __forceinline__ void slot_terms(const BB& b, int32_t s, int32_t a, int32_t* first, int32_t* count,
                                           int32_t* r, int32_t* rb, int32_t* L, int32_t* A) {
    const int32_t n = b.act[a];
    *first = b.nfirst[n]; *count = b.ncount[n];
    *r = s - *first;
    *rb = b.scan[s] - b.scan[*first];
    *L = b.aleft[a];
    *A = (*L < *count && b.flag[*first + *L]) ? *L + 1 : *L;
}

__global__ __launch_bounds__(BT) void k_bb_ranks(BB b) {
    const int32_t s = blockIdx.x * BT + threadIdx.x;
    if (s >= b.T) return;
    const int32_t a = b.slot_act[s];
    if (a < 0) return;
    int32_t first, count, r, rb, L, A;
    slot_terms(b, s, a, &first, &count, &r, &rb, &L, &A);
    const int32_t f = b.flag[s];
    if (r < A) { if (f) b.fr[first + rb] = r; }
    else if (!f) b.bl[first + (L - (r - rb) - 1)] = r;
}

__global__ __launch_bounds__(BT) void k_bb_scatter(BB b) {
    const int32_t s = blockIdx.x * BT + threadIdx.x;
    if (s >= b.T) return;
    const int32_t a = b.slot_act[s];
    if (a < 0) { b.idx2[s] = b.idx[s]; return; }
    int32_t first, count, r, rb, L, A;
    slot_terms(b, s, a, &first, &count, &r, &rb, &L, &A);
    const int32_t f = b.flag[s];
    int32_t dest;
    if (r < A) dest = !f ? r : (rb == 0 ? count - 1 : b.bl[first + rb - 1] - 1);
    else dest = f ? r - 1 : b.fr[first + (L - (r - rb) - 1)];
    if (dest < 0 || dest >= count) { b.counters[2] = 1; return; }
    b.idx2[first + dest] = b.idx[s];
}

// Split-node counts of subtrees, bottom-up over one level's nodes.
__global__ __launch_bounds__(BT) void k_bb_icount(BB b, int32_t lo, int32_t hi, int32_t* icnt) {
    const int32_t n = lo + blockIdx.x * BT + threadIdx.x;
    if (n >= hi) return;
    const int32_t c = b.nchild[n];
    icnt[n] = c < 0 ? 0 : 1 + icnt[c] + icnt[c + 1];
}

// Pre-order ranks and ids of one level's children, top-down.
__global__ __launch_bounds__(BT) void k_bb_ids(BB b, int32_t lo, int32_t hi, const int32_t* icnt, int32_t* rank, int32_t* id) {
    const int32_t n = lo + blockIdx.x * BT + threadIdx.x;
    if (n >= hi) return;
    const int32_t c = b.nchild[n];
    if (c < 0) return;
    rank[c] = rank[n] + 1;
    rank[c + 1] = rank[n] + 1 + icnt[c];
    id[c] = 1 + 2 * rank[n];
    id[c + 1] = 2 + 2 * rank[n];
}

__global__ __launch_bounds__(BT) void k_bb_emit(BB b, int32_t nn, const int32_t* id, const int64_t* node_base,
                                                int64_t total, Node* out) {
    const int32_t n = blockIdx.x * BT + threadIdx.x;
    if (n >= nn) return;
    const int32_t m = b.nmesh[n];
    if (b.ncount[n] == 0) return;                 // root of an empty mesh
    const int64_t at = node_base[m] + id[n], end = m + 1 < b.M ? node_base[m + 1] : total;
    if (id[n] < 0 || at >= end) { b.counters[2] = 1; return; }
    Node o;
    for (int k = 0; k < 3; ++k) { o.lo[k] = b.nbox[6 * (size_t)n + k]; o.hi[k] = b.nbox[6 * (size_t)n + 3 + k]; }
    const int32_t c = b.nchild[n];
    if (c >= 0) { o.a = (uint32_t)id[c]; o.b = 0u; }
    else { o.a = (uint32_t)(b.nfirst[n] - b.mesh_first[m]); o.b = (uint32_t)b.ncount[n]; }
    out[at] = o;
}

inline int32_t blocks_for(int64_t n) { return (int32_t)((n + BT - 1) / BT); }

// The widest launch of a build is one thread per bin word of a level's
// nodes: at most (T / (LEAF_MAX + 1) + 1) * BIN_WORDS threads for T up to
// BLAS_DEVICE_MAX_TRIS; its block count must fit the int32 grid size above,
// and every per-slot and per-node index (< 2T) an int32.
constexpr int64_t MAX_ACT = (int64_t)BLAS_DEVICE_MAX_TRIS / (LEAF_MAX + 1) + 1;
static_assert((MAX_ACT * BIN_WORDS + BT - 1) / BT < ((int64_t)1 << 31), "bin-word grid fits an int32 block count");
static_assert(2 * (int64_t)BLAS_DEVICE_MAX_TRIS < ((int64_t)1 << 31), "slot and node indices fit int32");

// Launch over n threads; nothing to do for n == 0 (a level whose splits all
// aborted adds no nodes).
template <typename K, typename... Args>
void launch_n(K kernel, int64_t n, hipStream_t st, Args... args) {
    if (n > 0) hipLaunchKernelGGL(kernel, dim3(blocks_for(n)), dim3(BT), 0, st, args...);
}

// Sub-buffers of one device allocation, 256-byte aligned.
struct Arena {
    size_t off = 0;
    char* base = nullptr;
    template <typename T>
    void take(T** p, size_t n) {
        *p = (T*)(base + off);
        off += (n * sizeof(T) + 255) & ~(size_t)255;
    }
};

}  // namespace

hipError_t device_build_blas(hipStream_t st, const float* pos_host, const int32_t* mesh_first, const int32_t* mesh_count,
                             int32_t M, int32_t T, std::vector<Node>* nodes, std::vector<int32_t>* order,
                             std::vector<int32_t>* used, std::vector<int32_t>* max_depth) {
    nodes->clear(); order->assign(T, 0); used->assign(M, 0); max_depth->assign(M, 0);
    if (M <= 0 || T <= 0) return hipSuccess;
    if (T > BLAS_DEVICE_MAX_TRIS) return hipErrorInvalidValue;
    char* arena = nullptr;
    hipError_t e = hipSuccess;
#define BB_CK(x) do { e = (x); if (e != hipSuccess) goto done; } while (0)
    {
        const int64_t node_cap = 2 * (int64_t)T + M;
        const int32_t act_cap = T / (LEAF_MAX + 1) + 1;
        const int32_t task_cap = T / TASK + act_cap + 1;
        std::vector<int64_t> base(M, 0);
        int64_t total_nodes = 0;
        for (int32_t m = 0; m < M; ++m) { base[m] = total_nodes; total_nodes += mesh_count[m] > 0 ? 2 * (int64_t)mesh_count[m] - 1 : 0; }
        std::vector<int2> root_tasks;
        for (int32_t m = 0; m < M; ++m)
            for (int32_t j = 0; j * TASK < mesh_count[m]; ++j) root_tasks.push_back(make_int2(m, j));
        size_t scan_bytes = 0;
        BB_CK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr, T + 1, st));
        BB b{};
        b.T = T; b.M = M;
        b.node_cap = (int32_t)node_cap; b.act_cap = act_cap; b.task_cap = task_cap;
        float* d_pos; int32_t* d_act[2]; int32_t *d_mfirst, *d_root_act, *icnt, *rank, *id; int64_t* d_base; Node* d_out;
        int2* d_root_tasks; void* d_scan_tmp;
        Arena A;
        for (int pass = 0; pass < 2; ++pass) {       // sizes, then pointers
            A.off = 0;
            A.take(&d_pos, 9 * (size_t)T); A.take(&b.cen, 3 * (size_t)T);
            A.take(&b.idx, T); A.take(&b.idx2, T); A.take(&b.slot_act, T);
            A.take(&b.flag, T + 1); A.take(&b.scan, T + 1); A.take(&b.fr, T); A.take(&b.bl, T);
            A.take(&b.nfirst, node_cap); A.take(&b.ncount, node_cap); A.take(&b.nchild, node_cap);
            A.take(&b.nmesh, node_cap); A.take(&b.ndepth, node_cap); A.take(&b.nnext, node_cap);
            A.take(&b.nkey, 6 * (size_t)node_cap); A.take(&b.nbox, 6 * (size_t)node_cap);
            A.take(&d_act[0], act_cap); A.take(&d_act[1], act_cap);
            A.take(&b.acb, 6 * (size_t)act_cap); A.take(&b.abin, (size_t)BIN_WORDS * act_cap);
            A.take(&b.aaxis, act_cap); A.take(&b.asplit, act_cap); A.take(&b.aleft, act_cap);
            A.take(&b.task, task_cap); A.take(&d_root_tasks, root_tasks.size());
            A.take(&b.counters, 4); A.take(&b.mesh_depth, M);
            A.take(&d_mfirst, M); A.take(&d_root_act, M);
            A.take(&icnt, node_cap); A.take(&rank, node_cap); A.take(&id, node_cap);
            A.take(&d_base, M); A.take(&d_out, total_nodes);
            A.take((char**)&d_scan_tmp, scan_bytes);
            if (pass == 0) { BB_CK(hipMalloc((void**)&arena, A.off)); A.base = arena; }
        }
        b.pos = d_pos;
        b.mesh_first = d_mfirst;

        // roots: node m for mesh m
        std::vector<int32_t> rfirst(M), rcount(M), rminus(M, -1), rzero(M, 0), rmesh(M), ract(M, -1), ract_list;
        std::vector<unsigned long long> rkey(6 * (size_t)M);
        for (int32_t m = 0; m < M; ++m) {
            rfirst[m] = mesh_first[m]; rcount[m] = mesh_count[m]; rmesh[m] = m;
            if (mesh_count[m] > LEAF_MAX) { ract[m] = (int32_t)ract_list.size(); ract_list.push_back(m); }
            for (int k = 0; k < 3; ++k) {
                rkey[6 * (size_t)m + k] = (0x80000000ull | 0x7F7FFFFFull) << 32 | LOW;   // key of FLT_MAX, initial position
                rkey[6 * (size_t)m + 3 + k] = (uint64_t)(~0xFF7FFFFFu) << 32;             // key of -FLT_MAX
            }
        }
        BB_CK(hipMemcpyAsync(d_pos, pos_host, 36 * (size_t)T, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(d_mfirst, mesh_first, 4 * (size_t)M, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(d_root_act, ract.data(), 4 * (size_t)M, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(b.nfirst, rfirst.data(), 4 * (size_t)M, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(b.ncount, rcount.data(), 4 * (size_t)M, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(b.nchild, rminus.data(), 4 * (size_t)M, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(b.nmesh, rmesh.data(), 4 * (size_t)M, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(b.ndepth, rzero.data(), 4 * (size_t)M, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(b.nnext, ract.data(), 4 * (size_t)M, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(b.nkey, rkey.data(), 48 * (size_t)M, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(b.mesh_depth, rzero.data(), 4 * (size_t)M, hipMemcpyHostToDevice, st));
        BB_CK(hipMemcpyAsync(d_base, base.data(), 8 * (size_t)M, hipMemcpyHostToDevice, st));
        if (!root_tasks.empty())
            BB_CK(hipMemcpyAsync(d_root_tasks, root_tasks.data(), sizeof(int2) * root_tasks.size(), hipMemcpyHostToDevice, st));
        if (!ract_list.empty())
            BB_CK(hipMemcpyAsync(d_act[0], ract_list.data(), 4 * ract_list.size(), hipMemcpyHostToDevice, st));
        int32_t host_counters[4] = {M, 0, 0, 0};
        BB_CK(hipMemcpyAsync(b.counters, host_counters, 16, hipMemcpyHostToDevice, st));
        BB_CK(hipMemsetAsync(d_out, 0, sizeof(Node) * (size_t)(total_nodes > 0 ? total_nodes : 1), st));
        BB_CK(hipMemsetAsync(rank, 0, 4 * (size_t)M, st));
        BB_CK(hipMemsetAsync(id, 0, 4 * (size_t)M, st));
        b.act = d_act[0]; b.act_next = d_act[1];
        launch_n(k_bb_init, T + 1, st, b, d_root_act);
        if (!root_tasks.empty())
            hipLaunchKernelGGL(k_bb_box, dim3((unsigned)root_tasks.size()), dim3(TW), 0, st, b, 0, (const int2*)d_root_tasks,
                               (int32_t)root_tasks.size());
        launch_n(k_bb_decode, M, st, b, 0, M);
        BB_CK(hipGetLastError());

        std::vector<int32_t> lvl = {0, M};           // level d holds nodes [lvl[d], lvl[d+1])
        int32_t nact = (int32_t)ract_list.size(), cur = 0;
        while (nact > 0) {
            if ((int64_t)lvl.size() > (int64_t)T + 2) { e = hipErrorUnknown; goto done; }   // depth is at most T
            b.act = d_act[cur]; b.act_next = d_act[cur ^ 1];
            const int32_t lo = lvl.back();
            const unsigned max_tasks = (unsigned)(T / TASK + nact + 1);
            launch_n(k_bb_level_init, (int64_t)nact * BIN_WORDS, st, b, nact);
            launch_n(k_bb_tasks, nact, st, b, nact);
            hipLaunchKernelGGL(k_bb_bounds, dim3(max_tasks), dim3(TW), 0, st, b);
            hipLaunchKernelGGL(k_bb_bins, dim3(max_tasks), dim3(TW), 0, st, b);
            launch_n(k_bb_split, nact, st, b, nact);
            launch_n(k_bb_flag, T, st, b);
            BB_CK(hipGetLastError());
            BB_CK(hipcub::DeviceScan::ExclusiveSum(d_scan_tmp, scan_bytes, b.flag, b.scan, T + 1, st));
            launch_n(k_bb_left, nact, st, b, nact);
            launch_n(k_bb_ranks, T, st, b);
            launch_n(k_bb_scatter, T, st, b);
            std::swap(b.idx, b.idx2);
            hipLaunchKernelGGL(k_bb_box, dim3(max_tasks), dim3(TW), 0, st, b, 1, (const int2*)b.task, 0);
            launch_n(k_bb_decode, 2 * (int64_t)nact, st, b, lo, -1);
            BB_CK(hipGetLastError());
            BB_CK(hipMemcpyAsync(host_counters, b.counters, 16, hipMemcpyDeviceToHost, st));
            BB_CK(hipStreamSynchronize(st));
            if (host_counters[2] || host_counters[0] > node_cap || host_counters[1] > act_cap) { e = hipErrorUnknown; goto done; }
            lvl.push_back(host_counters[0]);
            nact = host_counters[1];
            cur ^= 1;
        }
        // pre-order renumbering
        const int32_t nn = lvl.back();
        for (int d = (int)lvl.size() - 2; d >= 0; --d)
            launch_n(k_bb_icount, lvl[d + 1] - lvl[d], st, b, lvl[d], lvl[d + 1], icnt);
        for (int d = 0; d + 1 < (int)lvl.size(); ++d)
            launch_n(k_bb_ids, lvl[d + 1] - lvl[d], st, b, lvl[d], lvl[d + 1], (const int32_t*)icnt, rank, id);
        launch_n(k_bb_emit, nn, st, b, nn, (const int32_t*)id, (const int64_t*)d_base,
                           total_nodes, d_out);
        BB_CK(hipGetLastError());
        std::vector<int32_t> root_icnt(M);
        BB_CK(hipMemcpyAsync(root_icnt.data(), icnt, 4 * (size_t)M, hipMemcpyDeviceToHost, st));
        BB_CK(hipMemcpyAsync(max_depth->data(), b.mesh_depth, 4 * (size_t)M, hipMemcpyDeviceToHost, st));
        BB_CK(hipMemcpyAsync(host_counters, b.counters, 16, hipMemcpyDeviceToHost, st));
        BB_CK(hipStreamSynchronize(st));
        if (host_counters[2]) { e = hipErrorUnknown; goto done; }
        // only the allocated nodes travel back (about a fifth of the 2*count-1
        // capacity with leaves of up to 8 triangles); the tails stay zero
        nodes->resize((size_t)total_nodes);
        for (int32_t m = 0; m < M; ++m) {
            (*used)[m] = mesh_count[m] > 0 ? 1 + 2 * root_icnt[m] : 0;
            if (mesh_count[m] > 0 && (*used)[m] > 2 * mesh_count[m] - 1) { e = hipErrorUnknown; goto done; }
            if ((*used)[m] > 0)
                BB_CK(hipMemcpyAsync(nodes->data() + base[m], d_out + base[m], sizeof(Node) * (size_t)(*used)[m],
                                     hipMemcpyDeviceToHost, st));
        }
        BB_CK(hipMemcpyAsync(order->data(), b.idx, 4 * (size_t)T, hipMemcpyDeviceToHost, st));
        BB_CK(hipStreamSynchronize(st));
    }
done:
#undef BB_CK
    if (e != hipSuccess) (void)hipStreamSynchronize(st);
    if (arena) (void)hipFree(arena);
    return e;
}

}  // namespace kpt
