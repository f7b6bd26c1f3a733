// kpt_vec.h — float vector/matrix helpers for the gfx950 kernels.
// Each operator evaluates in the same association order as the reference's
// helper_math / Matrix4x4 code (vector_types_extension.cuh, matrix_maths.cuh)
// so that, compiled with -ffp-contract=off, results match the CPU oracle bit
// for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../common/kpt_dmath.h"

#define KD __host__ __device__ __forceinline__

namespace kpt {

struct v2 { float x, y; };
struct v3 { float x, y, z; };
struct v4 { float x, y, z, w; };

KD v3 mk3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
KD v4 mk4(float x, float y, float z, float w) { v4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
KD v4 mk4(v3 a, float w) { return mk4(a.x, a.y, a.z, w); }
KD v3 xyz(v4 a) { return mk3(a.x, a.y, a.z); }
KD v3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

KD v3 operator+(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
KD v3 operator-(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
KD v3 operator*(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
KD v3 operator/(v3 a, v3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
KD v3 operator+(v3 a, float b) { return mk3(a.x + b, a.y + b, a.z + b); }
KD v3 operator-(v3 a, float b) { return mk3(a.x - b, a.y - b, a.z - b); }
KD v3 operator*(v3 a, float b) { return mk3(a.x * b, a.y * b, a.z * b); }
KD v3 operator*(float b, v3 a) { return mk3(b * a.x, b * a.y, b * a.z); }
KD v3 operator/(v3 a, float b) { return mk3(a.x / b, a.y / b, a.z / b); }
KD v3 operator/(float b, v3 a) { return mk3(b / a.x, b / a.y, b / a.z); }
KD v3 operator-(float b, v3 a) { return mk3(b - a.x, b - a.y, b - a.z); }
KD v3 operator-(v3 a) { return mk3(-a.x, -a.y, -a.z); }
KD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KD float dot4(v4 a, v4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
KD v3 cross(v3 a, v3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
KD float sqr(float v) { return v * v; }
KD float pow5(float v) { return v * v * v * v * v; }
KD float length(v3 v) { return __builtin_sqrtf(dot(v, v)); }
KD float distance(v3 a, v3 b) { return length(a - b); }
KD v3 normalize(v3 v) { float inv = 1.0f / __builtin_sqrtf(dot(v, v)); return v * inv; }
KD v3 reflect(v3 i, v3 n) { return i - (2.0f * n) * dot(n, i); }
KD float fminf_(float a, float b) { return __builtin_fminf(a, b); }
KD float fmaxf_(float a, float b) { return __builtin_fmaxf(a, b); }
KD float clampf(float f, float a, float b) { return fmaxf_(a, fminf_(f, b)); }
KD float saturate(float v) { return clampf(v, 0.0f, 1.0f); }
KD v3 fmin3(v3 a, v3 b) { return mk3(fminf_(a.x, b.x), fminf_(a.y, b.y), fminf_(a.z, b.z)); }
KD v3 fmax3(v3 a, v3 b) { return mk3(fmaxf_(a.x, b.x), fmaxf_(a.y, b.y), fmaxf_(a.z, b.z)); }
KD v3 clamp3(v3 v, v3 a, v3 b) { return mk3(clampf(v.x, a.x, b.x), clampf(v.y, a.y, b.y), clampf(v.z, a.z, b.z)); }
KD bool is_zero(v3 v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f; }
KD float max_comp(v3 v) { return fmaxf_(v.x, fmaxf_(v.y, v.z)); }
KD float absdot(v3 a, v3 b) { return __builtin_fabsf(dot(a, b)); }

// Column-major 4x4 (matrix_maths.cuh:172-358); c[col*4 + row].
KD v4 mul4(const float* m, v4 v) {
    v4 r;
    r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
    r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
    r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
    r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
    return r;
}
// Only the x,y,z rows (the reference discards w in make_float3(M*v)).
KD v3 mul4_xyz(const float* m, v4 v) {
    return mk3(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w,
               m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w,
               m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w);
}

// 3x3 basis stored as three columns (Matrix3x3, matrix_maths.cuh:13-169).
struct m3 { v3 c0, c1, c2; };
KD v3 mul3(const m3& m, v3 v) {
    return mk3(m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z,
               m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
               m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z);
}
// transpose(m) * v
KD v3 mul3t(const m3& m, v3 v) {
    return mk3(m.c0.x * v.x + m.c0.y * v.y + m.c0.z * v.z,
               m.c1.x * v.x + m.c1.y * v.y + m.c1.z * v.z,
               m.c2.x * v.x + m.c2.y * v.y + m.c2.z * v.z);
}
// generateONBFrisvad, matrix_maths.cuh:375-396
KD m3 onb_frisvad(v3 n) {
    m3 r;
    if (n.z < -0.999805696f) {
        r.c0 = mk3(0.0f, -1.0f, 0.0f);
        r.c1 = mk3(-1.0f, 0.0f, 0.0f);
    } else {
        float a = 1.0f / (1.0f + n.z);
        float b = -n.x * n.y * a;
        r.c0 = mk3(1.0f - n.x * n.x * a, b, -n.x);
        r.c1 = mk3(b, 1.0f - n.y * n.y * a, -n.y);
    }
    r.c2 = n;
    return r;
}

}  // namespace kpt
