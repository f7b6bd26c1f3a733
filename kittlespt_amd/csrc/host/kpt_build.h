// kpt_build.h — host acceleration-structure builders.
#pragma once
#include <cstdint>
#include <vector>

#include "../common/kpt_device_types.h"

namespace kpt {

struct Aabb {
    float lo[3];
    float hi[3];
};

// One bottom-level tree over the triangles [first, first+count) of `pos`
// (9 floats per triangle).  Nodes use local indices (root = 0, children
// allocated in (left, left+1) pairs starting at 1); `order` is the slot ->
// triangle permutation (the reference's tri_index slice).
struct BlasTree {
    std::vector<Node> nodes;       // 2*count-1 entries, unused tail zeroed
    std::vector<int32_t> order;    // count entries, global triangle ids
    int32_t used = 0;              // nodes actually allocated
    int32_t max_depth = 0;
    Aabb bounds;                   // object-space bounds of the root
};

// BLASBuilder (as_builder.cuh:11-295): centroid = (p0+p1+p2)*0.3333f, 8 bins
// per axis over the centroid extent, SAH sweep, leaves of <= 8 triangles,
// in-place two-pointer partition, abort on an empty side, pre-order pair
// allocation.
BlasTree build_blas(const float* pos, int32_t first, int32_t count);

// TLASBuilder (as_builder.cuh:297-392): agglomerative clustering of instance
// world boxes by smallest merged half-area (findBestMatch), interior children
// packed 16|16 bits, the root copied to node 0.  Returns the nodes; `root` is
// -1 for an empty scene.  max 65535 nodes (16-bit child fields).
struct TlasTree {
    std::vector<Node> nodes;
    int32_t root = -1;
    int32_t max_depth = 0;
};
TlasTree build_tlas(const std::vector<Aabb>& inst_bounds);

// World box of an instance from its object box's 8 corners
// (BLAS2::setTransform, blas.cuh:190-200).
Aabb transform_bounds(const Aabb& b, const float model[16]);

}  // namespace kpt
