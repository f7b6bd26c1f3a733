// kpt_tlas.hip — the TLAS rebuilt on the device after instance edits.
//
// k_tlas_build is TLASBuilder::buildTLAS (as_builder.cuh:312-392) step for
// step: the same agglomerative clustering over the instances' world boxes
// (a nearest-neighbour chain: find A's best match B, B's best match C, merge
// when C == A, else walk on), the same node numbering (leaves 1..N in
// instance order, merged nodes from N+1, node 0 a copy of the root) and the
// same arithmetic (findBestMatch's half area of the merged box, first minimum
// in list order; merged boxes by fminf/fmaxf).  The loop is sequential; each
// best-match scan is a reduction over the live list by one workgroup, with
// the list's boxes in LDS.  The result equals the host builder's node for
// node.  The walk's ~2.5 N dependent scans make it slower than the host
// builder on 1,000 instances (about 4 ms against about 1.5 ms per rebuild,
// with 64, 128 or 256 threads), so the host builds by default; this path
// keeps the rebuild on the stream for callers that must not wait on it.  k_tlas_layout and k_tlas_blocks then derive the traversal's pair
// layout and two-level node blocks from it, as the host does.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../common/kpt_device_types.h"
#include "kpt_launch.h"

namespace kpt {
namespace {

constexpr int TB_THREADS = 256;
constexpr int TB_WAVES = TB_THREADS / 64;

// fminf / fmaxf as the host's C library evaluates them on the merged boxes:
// the second operand on equality (so +0/-0 ties keep its sign), the other
// operand when one is NaN.
__device__ __forceinline__ float fmin_host(float x, float y) { return (x < y || y != y) ? x : y; }
__device__ __forceinline__ float fmax_host(float x, float y) { return (x > y || y != y) ? x : y; }

struct TlasLds {
    float lo[3][TLAS_DEVICE_MAX];
    float hi[3][TLAS_DEVICE_MAX];
    int32_t ids[TLAS_DEVICE_MAX];
    float part_sa[TB_WAVES];
    int32_t part_idx[TB_WAVES];
};

// findBestMatch(list, n, X): the candidate B != X with the smallest half area
// of the box around X and B, the first such B in list order; -1 when no
// candidate's area is below infinity.  Called by every thread of the block.
__device__ int best_match(TlasLds& s, int n, int X) {
    const float ax0 = s.lo[0][X], ay0 = s.lo[1][X], az0 = s.lo[2][X];
    const float ax1 = s.hi[0][X], ay1 = s.hi[1][X], az1 = s.hi[2][X];
    float bs = __builtin_inff();
    int bi = 0x7FFFFFFF;
    for (int j = (int)threadIdx.x; j < n; j += TB_THREADS) {
        if (j == X) continue;
        const float bx1 = s.hi[0][j], by1 = s.hi[1][j], bz1 = s.hi[2][j];
        const float bx0 = s.lo[0][j], by0 = s.lo[1][j], bz0 = s.lo[2][j];
        const float ex = (ax1 > bx1 ? ax1 : bx1) - (ax0 < bx0 ? ax0 : bx0);
        const float ey = (ay1 > by1 ? ay1 : by1) - (ay0 < by0 ? ay0 : by0);
        const float ez = (az1 > bz1 ? az1 : bz1) - (az0 < bz0 ? az0 : bz0);
        const float sa = ex * ey + ey * ez + ez * ex;
        if (sa < bs) { bs = sa; bi = j; }   // j grows: the thread keeps its first minimum
    }
    // (area, index) minimum: a smaller area, or an equal area at a smaller index
    for (int off = 32; off > 0; off >>= 1) {
        const float os = __shfl_xor(bs, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (os < bs || (os == bs && oi < bi)) { bs = os; bi = oi; }
    }
    const int w = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { s.part_sa[w] = bs; s.part_idx[w] = bi; }
    __syncthreads();
    bs = s.part_sa[0];
    bi = s.part_idx[0];
    for (int k = 1; k < TB_WAVES; ++k) {
        const float os = s.part_sa[k];
        const int oi = s.part_idx[k];
        if (os < bs || (os == bs && oi < bi)) { bs = os; bi = oi; }
    }
    __syncthreads();   // the partials are rewritten by the next call
    return bi == 0x7FFFFFFF ? -1 : bi;
}

// One workgroup.  bounds: N boxes (lo xyz, hi xyz); nodes: 2N entries in the
// reference's numbering.  status[0] = nodes written, or -1 when a best match
// does not exist or the walk exceeds its step bound (the host checks the
// boxes first; neither happens for finite boxes).
__global__ __launch_bounds__(TB_THREADS) void k_tlas_build(const float* bounds, int N, Node* nodes, int32_t* status) {
    __shared__ TlasLds s;
    for (int i = (int)threadIdx.x; i < N; i += TB_THREADS) {
        Node n;
        for (int k = 0; k < 3; ++k) {
            n.lo[k] = bounds[6 * i + k];
            n.hi[k] = bounds[6 * i + 3 + k];
            s.lo[k][i] = n.lo[k];
            s.hi[k][i] = n.hi[k];
        }
        n.a = 0u;              // leaf marker
        n.b = (uint32_t)i;     // instance
        nodes[1 + i] = n;
        s.ids[i] = 1 + i;
    }
    __syncthreads();
    int left = N, ptr = N + 1;
    int A = 0, B = best_match(s, left, A);
    bool ok = true;
    // The walk merges N - 1 times and moves on a few times per merge; the cap
    // (never reached by a terminating walk of this size) keeps a degenerate
    // input from running on without end.
    const int max_steps = 64 * N + 1024;
    for (int steps = 0; left > 1; ++steps) {
        if (B < 0 || steps > max_steps) { ok = false; break; }
        const int C = best_match(s, left, B);
        if (A == C) {
            if (threadIdx.x == 0) {
                const int na = s.ids[A], nb = s.ids[B];
                Node m;
                for (int k = 0; k < 3; ++k) {
                    m.lo[k] = fmin_host(s.lo[k][A], s.lo[k][B]);
                    m.hi[k] = fmax_host(s.hi[k][A], s.hi[k][B]);
                }
                m.a = (uint32_t)na + ((uint32_t)nb << 16);
                m.b = 0xFFFFFFFFu;
                nodes[ptr] = m;
                // the list update in the reference's order (A may be the last entry)
                s.ids[A] = ptr;
                s.ids[B] = s.ids[left - 1];
                for (int k = 0; k < 3; ++k) {
                    s.lo[k][A] = m.lo[k]; s.hi[k][A] = m.hi[k];
                    s.lo[k][B] = s.lo[k][left - 1]; s.hi[k][B] = s.hi[k][left - 1];
                }
            }
            __syncthreads();
            ++ptr;
            --left;
            B = best_match(s, left, A);
        } else {
            A = B;
            B = C;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (ok && N > 0) {
            nodes[0] = nodes[s.ids[A]];
            status[0] = ptr;
        } else {
            status[0] = -1;
        }
    }
}

__global__ void k_tlas_fill(int32_t* p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0;
}

// Pair layout: the children of the merged node with reference index i (i > N)
// go to the pair q = i - N, nodes 2q and 2q + 1; an interior node's a becomes
// its own children's first index; the root (node 0) keeps its place; node 1
// is unused.  One thread per merged node, one for the root.
__device__ __forceinline__ Node relink(Node n, int N, uint32_t self) {
    if (n.a != 0u) n.a = 2u * (self - (uint32_t)N);
    return n;
}

// Nothing is written unless the build completed (status[0] == 2N): the
// previous layout and blocks then stay in place.
__global__ void k_tlas_layout(const Node* ref, int N, Node* out, const int32_t* status) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (status[0] != 2 * N) return;
    if (t == 0) {
        Node root = ref[0];
        // the root is the last merged node, 2N - 1 (or leaf 1 when N == 1)
        if (root.a != 0u) root.a = 2u * (uint32_t)(N - 1);
        out[0] = root;
        out[1] = Node{};
    }
    const int i = N + 1 + t;
    if (t >= N - 1 || i >= 2 * N) return;
    const uint32_t q = (uint32_t)(i - N);
    const uint32_t a = ref[i].a;
    const uint32_t na = a & 0xFFFFu, nb = a >> 16;
    if (na >= (uint32_t)(2 * N) || nb >= (uint32_t)(2 * N)) return;
    out[2 * q] = relink(ref[na], N, na);
    out[2 * q + 1] = relink(ref[nb], N, nb);
}

// NodeBlock q: nodes 2q, 2q + 1 and the children of each interior one; block
// 0 and block N stay empty, as the host's make_blocks leaves them.
__global__ void k_tlas_blocks(const Node* pl, int N, NodeBlock* blocks, const int32_t* status) {
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q > N || status[0] != 2 * N) return;
    NodeBlock b{};
    if (q >= 1 && q < N) {
        b.c[0] = pl[2 * q];
        b.c[1] = pl[2 * q + 1];
        for (int k = 0; k < 2; ++k) {
            const Node& ch = b.c[k];
            if (ch.a != 0u && ch.a + 1u < (uint32_t)(2 * N)) {
                b.g[2 * k] = pl[ch.a];
                b.g[2 * k + 1] = pl[ch.a + 1];
            }
        }
    }
    blocks[q] = b;
}

}  // namespace

hipError_t launch_tlas_build(hipStream_t s, const float* bounds, int32_t n, Node* ref_nodes, Node* pair_nodes,
                             NodeBlock* blocks, int32_t* status) {
    if (n <= 0 || n > TLAS_DEVICE_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_tlas_build, dim3(1), dim3(TB_THREADS), 0, s, bounds, (int)n, ref_nodes, status);
    const unsigned g1 = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_tlas_layout, dim3(g1), dim3(256), 0, s, ref_nodes, (int)n, pair_nodes, status);
    const unsigned g2 = (unsigned)((n + 1 + 255) / 256);
    hipLaunchKernelGGL(k_tlas_blocks, dim3(g2), dim3(256), 0, s, pair_nodes, (int)n, blocks, status);
    return hipGetLastError();
}

}  // namespace kpt

namespace kpt {
namespace {

// Refit: the topology of the last build kept, every box recomputed from the
// instances' current boxes (leaves) up to the root, in parallel: one thread
// per leaf walks up through the parents, and at each parent the second of
// its two children to arrive (per-node arrival counters, cleared by the
// caller) computes its box, as the build merged it: fminf/fmaxf of the first
// child's box with the second's.  Node 0 (the root's copy) takes the root's
// box at the end.  status[0] = 2N afterwards, so the layout kernels run.
__global__ void k_tlas_refit(const float* bounds, int N, Node* nodes, const int32_t* parent, int32_t* arrivals,
                             int32_t* status) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= N) return;
    Node& leaf = nodes[1 + i];
    for (int k = 0; k < 3; ++k) {
        leaf.lo[k] = bounds[6 * i + k];
        leaf.hi[k] = bounds[6 * i + 3 + k];
    }
    int cur = 1 + i;
    for (;;) {
        const int p = parent[cur];
        if (p <= 0) break;                                // the root (or a single leaf)
        __threadfence();                                  // this node's box before the arrival
        if (atomicAdd(&arrivals[p], 1) == 0) return;      // the sibling finishes the parent
        __threadfence();
        // the children's boxes were written by other threads, maybe on other
        // compute units: read them past the (non-coherent) L1
        const volatile Node* vn = nodes;
        const uint32_t a = vn[p].a;
        const uint32_t na = a & 0xFFFFu, nb = a >> 16;
        Node& n = nodes[p];
        for (int k = 0; k < 3; ++k) {
            n.lo[k] = fmin_host(vn[na].lo[k], vn[nb].lo[k]);
            n.hi[k] = fmax_host(vn[na].hi[k], vn[nb].hi[k]);
        }
        cur = p;
    }
    // only the thread that completed the root gets here (or the single leaf)
    __threadfence();
    const volatile Node* vn = nodes;
    Node root;
    for (int k = 0; k < 3; ++k) { root.lo[k] = vn[cur].lo[k]; root.hi[k] = vn[cur].hi[k]; }
    root.a = vn[cur].a;
    root.b = vn[cur].b;
    nodes[0] = root;
    status[0] = 2 * N;
}

}  // namespace

hipError_t launch_tlas_refit(hipStream_t s, const float* bounds, int32_t n, Node* ref_nodes, const int32_t* parent,
                             int32_t* arrivals, Node* pair_nodes, NodeBlock* blocks, int32_t* status) {
    if (n <= 0) return hipErrorInvalidValue;
    const unsigned g = (unsigned)((2 * n + 255) / 256);
    hipLaunchKernelGGL(k_tlas_fill, dim3(g), dim3(256), 0, s, arrivals, 2 * n);
    hipLaunchKernelGGL(k_tlas_refit, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, bounds, (int)n, ref_nodes,
                       parent, arrivals, status);
    const unsigned g1 = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_tlas_layout, dim3(g1), dim3(256), 0, s, ref_nodes, (int)n, pair_nodes, status);
    const unsigned g2 = (unsigned)((n + 1 + 255) / 256);
    hipLaunchKernelGGL(k_tlas_blocks, dim3(g2), dim3(256), 0, s, pair_nodes, (int)n, blocks, status);
    return hipGetLastError();
}

}  // namespace kpt
