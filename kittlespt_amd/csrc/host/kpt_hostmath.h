// kpt_hostmath.h — host-side matrix helpers of the renderer (column-major
// 4x4 as 16 floats, element (row r, column c) at m[4*c + r]).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace kpt {
namespace hm {

struct Mat4 { float m[16]; };

inline Mat4 identity() {
    Mat4 r;
    std::memset(r.m, 0, sizeof(r.m));
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}
inline Mat4 load(const float* p) { Mat4 r; std::memcpy(r.m, p, sizeof(r.m)); return r; }

// glm::perspectiveFovLH with the [-1, 1] depth range (ext/matrix_clip_space.inl):
// h = cos(fov/2) / sin(fov/2), w = h * height / width.
inline Mat4 perspective_fov_lh(float fov, float width, float height, float znear, float zfar) {
    Mat4 r;
    std::memset(r.m, 0, sizeof(r.m));
    const float h = std::cos(0.5f * fov) / std::sin(0.5f * fov);
    const float w = h * height / width;
    r.m[0] = w;
    r.m[5] = h;
    r.m[4 * 2 + 3] = 1.0f;
    r.m[4 * 2 + 2] = (zfar + znear) / (zfar - znear);
    r.m[4 * 3 + 2] = -(2.0f * zfar * znear) / (zfar - znear);
    return r;
}

// Matrix4x4::inverse (matrix_maths.cuh:304-354): 2x2 sub-factors of the last
// two columns, cofactor expansion along the first column, identity when
// |det| < 1e-6.  Used for the traversal transform (blas.cuh:191) and for every
// device-side inverse of the reference (interaction.cu:18, kernels.cu:174-179).
inline Mat4 ref_inverse(const Mat4& M) {
    auto c = [&](int col, int row) { return M.m[4 * col + row]; };
    float s00 = c(2,2) * c(3,3) - c(3,2) * c(2,3);
    float s01 = c(2,1) * c(3,3) - c(3,1) * c(2,3);
    float s02 = c(2,1) * c(3,2) - c(3,1) * c(2,2);
    float s03 = c(2,0) * c(3,3) - c(3,0) * c(2,3);
    float s04 = c(2,0) * c(3,2) - c(3,0) * c(2,2);
    float s05 = c(2,0) * c(3,1) - c(3,0) * c(2,1);
    float det = c(0,0) * (c(1,1) * s00 - c(1,2) * s01 + c(1,3) * s02) -
                c(0,1) * (c(1,0) * s00 - c(1,2) * s03 + c(1,3) * s04) +
                c(0,2) * (c(1,0) * s01 - c(1,1) * s03 + c(1,3) * s05) -
                c(0,3) * (c(1,0) * s02 - c(1,1) * s04 + c(1,2) * s05);
    if (std::fabs(det) < 1e-6f) return identity();
    float inv = 1.0f / det;
    Mat4 o;
    auto set = [&](int col, int row, float v) { o.m[4 * col + row] = v; };
    set(0,0, (c(1,1) * s00 - c(1,2) * s01 + c(1,3) * s02) * inv);
    set(0,1, -(c(0,1) * s00 - c(0,2) * s01 + c(0,3) * s02) * inv);
    set(0,2, (c(0,1) * (c(1,2) * c(3,3) - c(1,3) * c(3,2)) - c(0,2) * (c(1,1) * c(3,3) - c(1,3) * c(3,1)) + c(0,3) * (c(1,1) * c(3,2) - c(1,2) * c(3,1))) * inv);
    set(0,3, -(c(0,1) * (c(1,2) * c(2,3) - c(1,3) * c(2,2)) - c(0,2) * (c(1,1) * c(2,3) - c(1,3) * c(2,1)) + c(0,3) * (c(1,1) * c(2,2) - c(1,2) * c(2,1))) * inv);
    set(1,0, -(c(1,0) * s00 - c(1,2) * s03 + c(1,3) * s04) * inv);
    set(1,1, (c(0,0) * s00 - c(0,2) * s03 + c(0,3) * s04) * inv);
    set(1,2, -(c(0,0) * (c(1,2) * c(3,3) - c(1,3) * c(3,2)) - c(0,2) * (c(1,0) * c(3,3) - c(1,3) * c(3,0)) + c(0,3) * (c(1,0) * c(3,2) - c(1,2) * c(3,0))) * inv);
    set(1,3, (c(0,0) * (c(1,2) * c(2,3) - c(1,3) * c(2,2)) - c(0,2) * (c(1,0) * c(2,3) - c(1,3) * c(2,0)) + c(0,3) * (c(1,0) * c(2,2) - c(1,2) * c(2,0))) * inv);
    set(2,0, (c(1,0) * s01 - c(1,1) * s03 + c(1,3) * s05) * inv);
    set(2,1, -(c(0,0) * s01 - c(0,1) * s03 + c(0,3) * s05) * inv);
    set(2,2, (c(0,0) * (c(1,1) * c(3,3) - c(1,3) * c(3,1)) - c(0,1) * (c(1,0) * c(3,3) - c(1,3) * c(3,0)) + c(0,3) * (c(1,0) * c(3,1) - c(1,1) * c(3,0))) * inv);
    set(2,3, -(c(0,0) * (c(1,1) * c(2,3) - c(1,3) * c(2,1)) - c(0,1) * (c(1,0) * c(2,3) - c(1,3) * c(2,0)) + c(0,3) * (c(1,0) * c(2,1) - c(1,1) * c(2,0))) * inv);
    set(3,0, -(c(1,0) * s02 - c(1,1) * s04 + c(1,2) * s05) * inv);
    set(3,1, (c(0,0) * s02 - c(0,1) * s04 + c(0,2) * s05) * inv);
    set(3,2, -(c(0,0) * (c(1,1) * c(3,2) - c(1,2) * c(3,1)) - c(0,1) * (c(1,0) * c(3,2) - c(1,2) * c(3,0)) + c(0,2) * (c(1,0) * c(3,1) - c(1,1) * c(3,0))) * inv);
    set(3,3, (c(0,0) * (c(1,1) * c(2,2) - c(1,2) * c(2,1)) - c(0,1) * (c(1,0) * c(2,2) - c(1,2) * c(2,0)) + c(0,2) * (c(1,0) * c(2,1) - c(1,1) * c(2,0))) * inv);
    return o;
}

// glm::inverse for mat4 as the reference's host code calls it (renderer.cu:257,
// 346-347, 412, 427-429): the vendored glm 1.0.0 scalar path computes 18
// 2x2 cofactors, four signed cofactor columns, the determinant as
// (d.x + d.y) + (d.z + d.w) of the first column times the first cofactor row,
// and scales by 1/det.  No singularity guard.
inline Mat4 glm_inverse(const Mat4& M) {
    auto m = [&](int col, int row) { return M.m[4 * col + row]; };
    float k00 = m(2,2) * m(3,3) - m(3,2) * m(2,3), k02 = m(1,2) * m(3,3) - m(3,2) * m(1,3), k03 = m(1,2) * m(2,3) - m(2,2) * m(1,3);
    float k04 = m(2,1) * m(3,3) - m(3,1) * m(2,3), k06 = m(1,1) * m(3,3) - m(3,1) * m(1,3), k07 = m(1,1) * m(2,3) - m(2,1) * m(1,3);
    float k08 = m(2,1) * m(3,2) - m(3,1) * m(2,2), k10 = m(1,1) * m(3,2) - m(3,1) * m(1,2), k11 = m(1,1) * m(2,2) - m(2,1) * m(1,2);
    float k12 = m(2,0) * m(3,3) - m(3,0) * m(2,3), k14 = m(1,0) * m(3,3) - m(3,0) * m(1,3), k15 = m(1,0) * m(2,3) - m(2,0) * m(1,3);
    float k16 = m(2,0) * m(3,2) - m(3,0) * m(2,2), k18 = m(1,0) * m(3,2) - m(3,0) * m(1,2), k19 = m(1,0) * m(2,2) - m(2,0) * m(1,2);
    float k20 = m(2,0) * m(3,1) - m(3,0) * m(2,1), k22 = m(1,0) * m(3,1) - m(3,0) * m(1,1), k23 = m(1,0) * m(2,1) - m(2,0) * m(1,1);
    const float f0[4] = {k00, k00, k02, k03}, f1[4] = {k04, k04, k06, k07}, f2[4] = {k08, k08, k10, k11};
    const float f3[4] = {k12, k12, k14, k15}, f4[4] = {k16, k16, k18, k19}, f5[4] = {k20, k20, k22, k23};
    const float a0[4] = {m(1,0), m(0,0), m(0,0), m(0,0)}, a1[4] = {m(1,1), m(0,1), m(0,1), m(0,1)};
    const float a2[4] = {m(1,2), m(0,2), m(0,2), m(0,2)}, a3[4] = {m(1,3), m(0,3), m(0,3), m(0,3)};
    float col[4][4];
    const float sa[4] = {1, -1, 1, -1}, sb[4] = {-1, 1, -1, 1};
    for (int i = 0; i < 4; ++i) {
        col[0][i] = ((a1[i] * f0[i] - a2[i] * f1[i]) + a3[i] * f2[i]) * sa[i];
        col[1][i] = ((a0[i] * f0[i] - a2[i] * f3[i]) + a3[i] * f4[i]) * sb[i];
        col[2][i] = ((a0[i] * f1[i] - a1[i] * f3[i]) + a3[i] * f5[i]) * sa[i];
        col[3][i] = ((a0[i] * f2[i] - a1[i] * f4[i]) + a2[i] * f5[i]) * sb[i];
    }
    float d0 = m(0,0) * col[0][0], d1 = m(0,1) * col[1][0], d2 = m(0,2) * col[2][0], d3 = m(0,3) * col[3][0];
    float det = (d0 + d1) + (d2 + d3);
    float one_over = 1.0f / det;
    Mat4 o;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) o.m[4 * c + r] = col[c][r] * one_over;
    return o;
}

// Matrix4x4::operator* (matrix_maths.cuh:269-283).
inline Mat4 mul(const Mat4& A, const Mat4& B) {
    Mat4 o;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            o.m[4 * c + r] = A.m[0 * 4 + r] * B.m[4 * c + 0] + A.m[1 * 4 + r] * B.m[4 * c + 1] +
                             A.m[2 * 4 + r] * B.m[4 * c + 2] + A.m[3 * 4 + r] * B.m[4 * c + 3];
    return o;
}

// Matrix4x4 * float4 (matrix_maths.cuh:258-267), xyz part.
inline void mul_point(const Mat4& M, const float v[4], float out[3]) {
    for (int r = 0; r < 3; ++r)
        out[r] = M.m[0 * 4 + r] * v[0] + M.m[1 * 4 + r] * v[1] + M.m[2 * 4 + r] * v[2] + M.m[3 * 4 + r] * v[3];
}

}  // namespace hm
}  // namespace kpt
