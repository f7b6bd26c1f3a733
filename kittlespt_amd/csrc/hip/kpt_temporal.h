// kpt_temporal.h — temporal accumulation of the trace kernels: reprojection of
// the history buffer through the motion vectors with a depth test
// (reprojectAccumulate / testReprojection / testReprojectedDepth,
// kernels.cu:196-338), shared by the megakernel and the wavefront finish pass.
#pragma once
#include "kpt_launch.h"
#include "kpt_shade.h"

namespace kpt {

constexpr float TEMPORAL_DEPTH_REJECT = 0.045f;   // kernels.cu:198
constexpr int MAX_ACCUMULATION_FRAMES = 16;       // kernels.cu:296

// DeviceTextureBuffer::textureReadNearest: truncate, clamp to the edge
KD int clamp_coord(int v, int n) { return v < 0 ? 0 : (v > n - 1 ? n - 1 : v); }

// testReprojection (kernels.cu:206-222): the surface moved by its previous
// model matrix must sit at the depth the previous G-buffer holds.
KD bool test_reprojection(const SceneView& S, const FrameConst& fc, const uint4* prev_gbuffer, int inst, v3 pos,
                          float ppx, float ppy) {
    const InstShade& is = S.inst_shade[inst];
    v3 local = mul4_xyz(is.curr_inv, mk4(pos, 1.0f));
    const int ix = clamp_coord((int)ppx, fc.width), iy = clamp_coord((int)ppy, fc.height);
    const float prev_depth = __builtin_bit_cast(float, prev_gbuffer[(size_t)iy * fc.width + ix].w);
    v3 prev_cam = mk3(fc.prev_cam_pos[0], fc.prev_cam_pos[1], fc.prev_cam_pos[2]);
    v3 prev_world = mul4_xyz(is.prev_model, mk4(local, 1.0f));
    v3 d = prev_cam - prev_world;
    const float reproj_depth = __builtin_sqrtf(dot(d, d));
    return __builtin_fabsf(reproj_depth - prev_depth) < (prev_depth * TEMPORAL_DEPTH_REJECT);
}

// reprojectAccumulate (kernels.cu:299-338): returns (colour, history length)
KD float4 reproject_accumulate(const SceneView& S, const FrameConst& fc, const FrameBuffers& fb, int px, int py,
                               int inst, v3 pos, v2 vel, v3 curr) {
    if (inst < 0) return make_float4(curr.x, curr.y, curr.z, 0.0f);
    const float ppx = (float)px - vel.x * (float)fc.width;
    const float ppy = (float)py - vel.y * (float)fc.height;
    if (ppx < 0 || ppx >= (float)fc.width || ppy < 0 || ppy >= (float)fc.height)
        return make_float4(curr.x, curr.y, curr.z, 0.0f);
    if (!test_reprojection(S, fc, fb.prev_gbuffer, inst, pos, ppx, ppy)) return make_float4(curr.x, curr.y, curr.z, 0.0f);
    // textureReadBilinear(prev, filter_alpha = false): alpha from the first tap
    const int dx = (int)ppx, dy = (int)ppy;
    const float ws = ppx - (float)dx, wt = ppy - (float)dy;
    auto tap = [&](int x, int y) {
        return fb.accum[(size_t)clamp_coord(y, fc.height) * fc.width + clamp_coord(x, fc.width)];
    };
    const float4 c0 = tap(dx, dy), c1 = tap(dx + 1, dy), c2 = tap(dx, dy + 1), c3 = tap(dx + 1, dy + 1);
    auto lerp1 = [](float a, float b, float t) { return a + t * (b - a); };
    v3 prev = mk3(lerp1(lerp1(c0.x, c1.x, ws), lerp1(c2.x, c3.x, ws), wt),
                  lerp1(lerp1(c0.y, c1.y, ws), lerp1(c2.y, c3.y, ws), wt),
                  lerp1(lerp1(c0.z, c1.z, ws), lerp1(c2.z, c3.z, ws), wt));
    const float history = c0.w;
    const float alpha = 1.0f / fminf_(history + 1, (float)MAX_ACCUMULATION_FRAMES);
    v3 o = mk3(lerp1(prev.x, curr.x, alpha), lerp1(prev.y, curr.y, alpha), lerp1(prev.z, curr.z, alpha));
    return make_float4(o.x, o.y, o.z, history + 1);
}

// End of a path for pixel (px, py): G/V-buffers, then static accumulation
// (Integrator::addSample, integrator.cuh:406-416) or temporal reprojection
// (kernels.cu:424-436) into the render buffer.
KD void write_pixel(const SceneView& S, const FrameConst& fc, const FrameBuffers& fb, int32_t frame, int64_t idx,
                    int px, int py, v3 L, v3 albedo, v3 normal, float depth, int inst, int prim, v3 pos) {
    fb.gbuffer[idx] = pack_gbuffer(albedo, normal, depth);
    v2 vel = velocity(S, fc, inst, pos);
    fb.vbuffer[idx] = make_float4(vel.x, vel.y, (float)prim, (float)inst);
    if (fc.temporal) {
        fb.render[idx] = reproject_accumulate(S, fc, fb, px, py, inst, pos, vel, L);
        return;
    }
    const float4 a = fb.accum[idx];
    v3 acc = mk3(a.x, a.y, a.z) + L;
    fb.accum[idx] = make_float4(acc.x, acc.y, acc.z, 1.0f);
    v3 est = acc / (float)(frame + 1);
    fb.render[idx] = make_float4(est.x, est.y, est.z, 1.0f);
}

}  // namespace kpt
