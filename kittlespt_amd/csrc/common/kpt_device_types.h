// kpt_device_types.h — HBM layouts shared by the host builders and the gfx950
// kernels.  Everything is plain 16-byte-aligned structs so kernels fetch with
// dwordx4 loads.
#pragma once
#include <stdint.h>

namespace kpt {

// BVH node, 32 bytes.  Same information as the reference's BVH2Node
// (blas.cuh:9-35: child/first, count, Bounds3f) and TLASNode
// (blas.cuh:209-238: Bounds3f, packed children, blas index) but split into
// two float4 so the slab test reads pmin and pmax with one load each.
//   BLAS interior: a = left child (global node index; right = a+1), b = 0
//   BLAS leaf:     a = first triangle slot (global), b = triangle count
//   TLAS:          a = left_right_index (16|16 bits, 0 for a leaf), b = instance
// Sibling pairs are allocated at even node indices so one 64-byte line holds
// both children of an interior node.
struct alignas(16) Node {
    float lo[3];
    uint32_t a;
    float hi[3];
    uint32_t b;
};
static_assert(sizeof(Node) == 32, "Node is 32 bytes");

// Two BVH levels in one 192-byte record, for the traversal kernels: the
// children pair of an interior node (c0, c1: copies of nodes 2q and 2q+1 of
// the level's pair layout, where q is the record index) and, for each
// interior child, its own children (g0, g1 for c0; g2, g3 for c1; zero
// otherwise).  BLAS records follow the BLAS node array (sibling pairs at even
// indices); TLAS records follow a renumbered copy of the TLAS in the same
// pair layout (root at 0, pairs from 2 on; interior a = children pair index,
// leaf a = 0, b = instance).
struct alignas(64) NodeBlock {
    Node c[2];
    Node g[4];
};
static_assert(sizeof(NodeBlock) == 192, "NodeBlock is 192 bytes");

// Triangle positions in BVH-leaf (slot) order: v0 and the two edges the
// Moller-Trumbore test needs (triangle.cuh:66-102 computes v1-v0 and v2-v0 per
// test; the host computes the same IEEE differences once).  v0.w holds the
// original triangle (primitive) id.
struct alignas(16) TriPos {
    float v0[3];
    int32_t prim;
    float e1[3];
    float pad0;
    float e2[3];
    float pad1;
};
static_assert(sizeof(TriPos) == 48, "TriPos is 48 bytes");

// Shading attributes in primitive order (Triangle, triangle.cuh:54-129).
struct alignas(16) TriAttr {
    float lgn[3];          // local geometric normal
    int32_t material;
    float uv0[2];
    float uv1[2];
    float uv2[2];
    int32_t light;         // -1 when not emissive
    int32_t inst_of_light; // instance used to place the light (first user)
};
static_assert(sizeof(TriAttr) == 48, "TriAttr is 48 bytes");

// Per-instance traversal record, 64 bytes: rows x, y, z of the reference's
// Mat4::inverse(model) (blas.cuh:190-192) so the object ray is three row dot
// products, plus the BLAS root node.
struct alignas(16) InstTrav {
    float r0[4];
    float r1[4];
    float r2[4];
    int32_t root;
    int32_t geom;
    uint32_t root_entry;   // traversal entry of the BLAS root (0xFFFFFFFF: empty)
    int32_t pad;
};
static_assert(sizeof(InstTrav) == 64, "InstTrav is 64 bytes");

// Per-instance shading record: column-major 4x4 matrices.
//   model     = Mat4::inverse(curr_inv_model)  (interaction.cu:18, light.cu:37)
//   curr_inv  = glm::inverse(model)            (TriangleMesh, renderer.cu:257)
//   prev_model= Mat4::inverse(prev_inv_model)  (kernels.cu:179)
struct alignas(16) InstShade {
    float model[16];
    float curr_inv[16];
    float prev_model[16];
};

// Material (material.cuh:21-75), 64 bytes.
struct alignas(16) Material {
    int32_t albedo_tex;
    float albedo[3];
    int32_t orm_tex;
    float metallic;
    float roughness;
    int32_t transmission_tex;
    float transmission;
    float ior;
    int32_t emission_tex;
    float emission_scale;
    float emission[3];
    int32_t normal_tex;
    float normal_scale;
    float pad[3];
};

// Texture (texture.cuh:21-65).
struct alignas(16) Texture {
    int32_t width;
    int32_t height;
    int32_t channels;
    int32_t bit_depth;
    int64_t offset;
    int64_t pad;
};

// Area light (light.cuh:68-95).
struct alignas(16) Light {
    float area;
    int32_t prim;
    int32_t inst;
    int32_t pad;
};

// Per-frame constants of the path tracer.
struct alignas(16) FrameConst {
    float inv_view[16];
    float inv_proj[16];
    float curr_vp[16];     // inverse(inv_proj) * inverse(inv_view), kernels.cu:174
    float prev_vp[16];     // kernels.cu:175
    float cam_pos[3];
    float sun_diam;
    float sun_dir[3];
    float sun_emission;
    float sun_radiance[3]; // sky sampleLi toward the sun: a per-frame constant
    float sun_radius;      // angularDiameterToPhysicalDiameter(...)/2
    float min_sun_cos;     // cos(diameter/2), LeSun
    int32_t width;
    int32_t height;
    int32_t max_depth;
    int32_t seed_mode;
    int32_t num_lights;
    int32_t num_instances;
    int32_t tlas_root;
    int32_t temporal;      // integrator_use_temporal_accumulation
    float prev_cam_pos[3]; // prev_inv_view column 3 (testReprojection, kernels.cu:215)
    float scene_lo[3];     // TLAS root box (spatial ray-sort keys)
    float scene_inv[3];    // 1 / extent per axis
};

// Scene pointers passed by value to kernels.
struct SceneView {
    const Node* blas_nodes;
    const Node* tlas_nodes;      // pair layout (see NodeBlock)
    const NodeBlock* blas_blocks;
    const NodeBlock* tlas_blocks;
    const TriPos* tris;
    const TriAttr* attrs;
    const InstTrav* inst_trav;
    const InstShade* inst_shade;
    const Material* materials;
    const Texture* textures;
    const uint8_t* pixels;
    const Light* lights;
    const float* light_verts;   // 9 floats per light: object-space vertices
    int64_t num_pixel_bytes;
    int32_t tlas_root;
    int32_t num_lights;
};

}  // namespace kpt
