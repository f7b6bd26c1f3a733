// kpt_oracle.cpp — CPU oracle: a restatement of the KittlesPT reference's hot
// path in IEEE semantics.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check
// and bench.py's cpu_baseline leg call it as the reference to compare with; the product library
// never links or loads it.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fno-fast-math).
// The float instantiation uses the deterministic transcendentals of
// kittlespt_amd/csrc/common/kpt_dmath.h (the GPU kernels use the same header;
// that header is checked against glibc independently in tests/test_oracle.py,
// test_dmath_*).
// The double instantiation uses libm and serves as the converged FP64 render.
//
// Every routine cites the reference file:line it restates.  Operation order
// follows the reference expression by expression (including float-vs-double
// literal promotion) so the float instantiation is a bit-level specification.
// Function-argument evaluation order, unspecified in C++, is defined here as
// left to right (e.g. IndependentSampler::get2D, samplers.cuh:179-182).
#include "kpt_oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "../kittlespt_amd/csrc/common/kpt_dmath.h"

namespace orc {

// ------------------------------------------------------------------ math
template <class R> struct Mth;
template <> struct Mth<float> {
    static float exp(float x) { return kpt::dm::expf(x); }
    static float exp2(float x) { return kpt::dm::exp2f(x); }
    static float log2(float x) { return kpt::dm::log2f(x); }
    static float pow(float x, float y) { return kpt::dm::powf(x, y); }
    static float sin(float x) { return kpt::dm::sinf(x); }
    static float cos(float x) { return kpt::dm::cosf(x); }
    static float tan(float x) { return kpt::dm::tanf(x); }
    static float acos(float x) { return kpt::dm::acosf(x); }
};
template <> struct Mth<double> {
    static double exp(double x) { return std::exp(x); }
    static double exp2(double x) { return std::exp2(x); }
    static double log2(double x) { return std::log2(x); }
    static double pow(double x, double y) { return std::pow(x, y); }
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double tan(double x) { return std::tan(x); }
    static double acos(double x) { return std::acos(x); }
};
// ---- op counting.  CF is a float that counts the arithmetic done on it,
// the algorithmic FP32 op count of a routine: +, -, *, / are arithmetic ops,
// square roots and transcendentals ops of their own kinds, min / max / abs /
// floor a fourth kind (comparisons, conversions and the integer RNG are not
// counted).  Counts go to the bucket of the current phase, so the oracle
// instantiated with CF tells how much of a path's work is traversal, sky and
// shading; the float and double instantiations ignore the phases.
enum OpKind { OP_ARITH, OP_SQRT, OP_TRANS, OP_MINMAX, OP_KINDS };
enum OpPhase { PH_TRAV, PH_SKY0, PH_SKY, PH_SKYSUN, PH_SHADE0, PH_SHADE, PH_OTHER, PH_COUNT };
struct OpCounts { uint64_t n[PH_COUNT][OP_KINDS]; };
thread_local OpCounts* g_ops = nullptr;
thread_local int g_phase = PH_OTHER;
inline void count_op(int k) { if (g_ops) g_ops->n[g_phase][k]++; }
struct CF {
    float v = 0.0f;
    CF() = default;
    CF(float x) : v(x) {}
    CF(double x) : v((float)x) {}
    CF(int x) : v((float)x) {}
    CF(uint32_t x) : v((float)x) {}
    CF(int64_t x) : v((float)x) {}
    explicit operator float() const { return v; }
    explicit operator double() const { return (double)v; }
    explicit operator int() const { return (int)v; }
    explicit operator uint32_t() const { return (uint32_t)v; }
    CF& operator+=(CF b) { count_op(OP_ARITH); v += b.v; return *this; }
    CF& operator-=(CF b) { count_op(OP_ARITH); v -= b.v; return *this; }
    CF& operator*=(CF b) { count_op(OP_ARITH); v *= b.v; return *this; }
    CF& operator/=(CF b) { count_op(OP_ARITH); v /= b.v; return *this; }
};
inline CF operator+(CF a, CF b) { count_op(OP_ARITH); return CF(a.v + b.v); }
inline CF operator-(CF a, CF b) { count_op(OP_ARITH); return CF(a.v - b.v); }
inline CF operator*(CF a, CF b) { count_op(OP_ARITH); return CF(a.v * b.v); }
inline CF operator/(CF a, CF b) { count_op(OP_ARITH); return CF(a.v / b.v); }
inline CF operator-(CF a) { return CF(-a.v); }
inline bool operator<(CF a, CF b) { return a.v < b.v; }
inline bool operator>(CF a, CF b) { return a.v > b.v; }
inline bool operator<=(CF a, CF b) { return a.v <= b.v; }
inline bool operator>=(CF a, CF b) { return a.v >= b.v; }
inline bool operator==(CF a, CF b) { return a.v == b.v; }
inline bool operator!=(CF a, CF b) { return a.v != b.v; }
template <> struct Mth<CF> {
    static CF exp(CF x) { count_op(OP_TRANS); return CF(kpt::dm::expf(x.v)); }
    static CF exp2(CF x) { count_op(OP_TRANS); return CF(kpt::dm::exp2f(x.v)); }
    static CF log2(CF x) { count_op(OP_TRANS); return CF(kpt::dm::log2f(x.v)); }
    static CF pow(CF x, CF y) { count_op(OP_TRANS); return CF(kpt::dm::powf(x.v, y.v)); }
    static CF sin(CF x) { count_op(OP_TRANS); return CF(kpt::dm::sinf(x.v)); }
    static CF cos(CF x) { count_op(OP_TRANS); return CF(kpt::dm::cosf(x.v)); }
    static CF tan(CF x) { count_op(OP_TRANS); return CF(kpt::dm::tanf(x.v)); }
    static CF acos(CF x) { count_op(OP_TRANS); return CF(kpt::dm::acosf(x.v)); }
};
// The phase of the work that follows (CF only); PhaseScope restores the
// previous one when it ends.
template <class R> inline void SetPhase(int) {}
template <> inline void SetPhase<CF>(int p) { g_phase = p; }
template <class R> struct PhaseScope { explicit PhaseScope(int) {} };
template <> struct PhaseScope<CF> {
    int old;
    explicit PhaseScope(int p) : old(g_phase) { g_phase = p; }
    ~PhaseScope() { g_phase = old; }
};

template <class R> R Sqrt(R x) { return std::sqrt(x); }
template <class R> R Fabs(R x) { return std::fabs(x); }
template <class R> R Fmin(R a, R b) { return std::fmin(a, b); }
template <class R> R Fmax(R a, R b) { return std::fmax(a, b); }
template <class R> R Floor(R x) { return std::floor(x); }
template <class R> R Inf() { return std::numeric_limits<R>::infinity(); }
template <> inline CF Sqrt<CF>(CF x) { count_op(OP_SQRT); return CF(std::sqrt(x.v)); }
template <> inline CF Fabs<CF>(CF x) { count_op(OP_MINMAX); return CF(std::fabs(x.v)); }
template <> inline CF Fmin<CF>(CF a, CF b) { count_op(OP_MINMAX); return CF(std::fmin(a.v, b.v)); }
template <> inline CF Fmax<CF>(CF a, CF b) { count_op(OP_MINMAX); return CF(std::fmax(a.v, b.v)); }
template <> inline CF Floor<CF>(CF x) { count_op(OP_MINMAX); return CF(std::floor(x.v)); }
template <> inline CF Inf<CF>() { return CF(std::numeric_limits<float>::infinity()); }
template <class R> bool IsNan(R x) { return x != x; }
template <class R> bool IsInf(R x) { return std::fabs((double)x) == (double)Inf<R>(); }
// vector_types_extension.cuh:1168-1171
template <class R> R Clamp(R f, R a, R b) { return Fmax(a, Fmin(f, b)); }
template <class R> R Saturate(R v) { return Clamp<R>(v, R(0), R(1)); }   // linear_algebra.cuh:65-67
template <class R> R Sqr(R v) { return v * v; }                          // linear_algebra.cuh:27-30
template <class R> R Pow5(R v) { return v * v * v * v * v; }             // linear_algebra.cuh:96-98

template <class R> struct V2 { R x = 0, y = 0; };
template <class R> struct V3 {
    R x = 0, y = 0, z = 0;
    V3() = default;
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
};
template <class R> struct V4 {
    R x = 0, y = 0, z = 0, w = 0;
    V4() = default;
    V4(R a, R b, R c, R d) : x(a), y(b), z(c), w(d) {}
    V4(V3<R> v, R d) : x(v.x), y(v.y), z(v.z), w(d) {}
};
template <class R> V3<R> xyz(V4<R> v) { return V3<R>(v.x, v.y, v.z); }
template <class R> V3<R> operator+(V3<R> a, V3<R> b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> V3<R> operator-(V3<R> a, V3<R> b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> V3<R> operator*(V3<R> a, V3<R> b) { return V3<R>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> V3<R> operator/(V3<R> a, V3<R> b) { return V3<R>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class R> V3<R> operator+(V3<R> a, R b) { return V3<R>(a.x + b, a.y + b, a.z + b); }
template <class R> V3<R> operator-(V3<R> a, R b) { return V3<R>(a.x - b, a.y - b, a.z - b); }
template <class R> V3<R> operator*(V3<R> a, R b) { return V3<R>(a.x * b, a.y * b, a.z * b); }
template <class R> V3<R> operator*(R b, V3<R> a) { return V3<R>(b * a.x, b * a.y, b * a.z); }
template <class R> V3<R> operator/(V3<R> a, R b) { return V3<R>(a.x / b, a.y / b, a.z / b); }
template <class R> V3<R> operator/(R b, V3<R> a) { return V3<R>(b / a.x, b / a.y, b / a.z); }
template <class R> V3<R> operator-(R b, V3<R> a) { return V3<R>(b - a.x, b - a.y, b - a.z); }
template <class R> V3<R> operator-(V3<R> a) { return V3<R>(-a.x, -a.y, -a.z); }
template <class R> R Dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> R Dot4(V4<R> a, V4<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
template <class R> V3<R> Cross(V3<R> a, V3<R> b) {
    return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> R Length(V3<R> v) { return Sqrt(Dot(v, v)); }
// normalize = v * rsqrtf(dot(v,v)) with host rsqrtf = 1/sqrtf
// (vector_types_extension.cuh:1325-1329, 78-81)
template <class R> V3<R> Normalize(V3<R> v) { R inv = R(1) / Sqrt(Dot(v, v)); return v * inv; }
// vector_types_extension.cuh:1427-1430: i - 2.0f * n * dot(n, i)
template <class R> V3<R> Reflect(V3<R> i, V3<R> n) { return i - (R(2) * n) * Dot(n, i); }
template <class R> V3<R> Fmin3(V3<R> a, V3<R> b) { return V3<R>(Fmin(a.x, b.x), Fmin(a.y, b.y), Fmin(a.z, b.z)); }
template <class R> V3<R> Fmax3(V3<R> a, V3<R> b) { return V3<R>(Fmax(a.x, b.x), Fmax(a.y, b.y), Fmax(a.z, b.z)); }
template <class R> V3<R> Clamp3(V3<R> v, V3<R> a, V3<R> b) {
    return V3<R>(Clamp(v.x, a.x, b.x), Clamp(v.y, a.y, b.y), Clamp(v.z, a.z, b.z));
}
template <class R> bool IsZero(V3<R> v) { return v.x == R(0) && v.y == R(0) && v.z == R(0); }  // color.cuh:145-148
template <class R> R MaxComp(V3<R> v) { return Fmax(v.x, Fmax(v.y, v.z)); }                     // color.cuh:304-307
template <class R> R AbsDot(V3<R> a, V3<R> b) { return Fabs(Dot(a, b)); }                       // vector_maths.cuh:11-14
template <class R> R Distance(V3<R> a, V3<R> b) { return Length(a - b); }
template <class R> V3<R> cvt(V3<float> v) { return V3<R>(R(v.x), R(v.y), R(v.z)); }

// Column-major 4x4 / 3x3 (matrix_maths.cuh:13-358)
template <class R> struct M4 {
    V4<R> c[4];
    M4() { c[0] = V4<R>(1, 0, 0, 0); c[1] = V4<R>(0, 1, 0, 0); c[2] = V4<R>(0, 0, 1, 0); c[3] = V4<R>(0, 0, 0, 1); }
};
template <class R> V4<R> Mul(const M4<R>& m, V4<R> v) {  // matrix_maths.cuh:258-267
    V4<R> r;
    r.x = m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z + m.c[3].x * v.w;
    r.y = m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z + m.c[3].y * v.w;
    r.z = m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z + m.c[3].z * v.w;
    r.w = m.c[0].w * v.x + m.c[1].w * v.y + m.c[2].w * v.z + m.c[3].w * v.w;
    return r;
}
template <class R> M4<R> Transpose(const M4<R>& m) {
    M4<R> t;
    t.c[0] = V4<R>(m.c[0].x, m.c[1].x, m.c[2].x, m.c[3].x);
    t.c[1] = V4<R>(m.c[0].y, m.c[1].y, m.c[2].y, m.c[3].y);
    t.c[2] = V4<R>(m.c[0].z, m.c[1].z, m.c[2].z, m.c[3].z);
    t.c[3] = V4<R>(m.c[0].w, m.c[1].w, m.c[2].w, m.c[3].w);
    return t;
}
template <class R> M4<R> MulMM(const M4<R>& a, const M4<R>& b) {  // matrix_maths.cuh:269-283
    M4<R> res;
    M4<R> m1 = Transpose(a);
    for (int row = 0; row < 4; ++row) {
        res.c[row].x = Dot4(m1.c[row], b.c[0]);
        res.c[row].y = Dot4(m1.c[row], b.c[1]);
        res.c[row].z = Dot4(m1.c[row], b.c[2]);
        res.c[row].w = Dot4(m1.c[row], b.c[3]);
    }
    return Transpose(res);
}
// Matrix4x4::inverse, matrix_maths.cuh:304-354
template <class R> M4<R> Inverse(const M4<R>& m) {
    const V4<R>& c0 = m.c[0]; const V4<R>& c1 = m.c[1]; const V4<R>& c2 = m.c[2]; const V4<R>& c3 = m.c[3];
    R s00 = c2.z * c3.w - c3.z * c2.w;
    R s01 = c2.y * c3.w - c3.y * c2.w;
    R s02 = c2.y * c3.z - c3.y * c2.z;
    R s03 = c2.x * c3.w - c3.x * c2.w;
    R s04 = c2.x * c3.z - c3.x * c2.z;
    R s05 = c2.x * c3.y - c3.x * c2.y;
    R det = c0.x * (c1.y * s00 - c1.z * s01 + c1.w * s02) -
            c0.y * (c1.x * s00 - c1.z * s03 + c1.w * s04) +
            c0.z * (c1.x * s01 - c1.y * s03 + c1.w * s05) -
            c0.w * (c1.x * s02 - c1.y * s04 + c1.z * s05);
    if (Fabs(det) < R(1e-6f)) return M4<R>();
    R inv = R(1) / det;
    M4<R> o;
    o.c[0].x = (c1.y * s00 - c1.z * s01 + c1.w * s02) * inv;
    o.c[0].y = -(c0.y * s00 - c0.z * s01 + c0.w * s02) * inv;
    o.c[0].z = (c0.y * (c1.z * c3.w - c1.w * c3.z) - c0.z * (c1.y * c3.w - c1.w * c3.y) + c0.w * (c1.y * c3.z - c1.z * c3.y)) * inv;
    o.c[0].w = -(c0.y * (c1.z * c2.w - c1.w * c2.z) - c0.z * (c1.y * c2.w - c1.w * c2.y) + c0.w * (c1.y * c2.z - c1.z * c2.y)) * inv;
    o.c[1].x = -(c1.x * s00 - c1.z * s03 + c1.w * s04) * inv;
    o.c[1].y = (c0.x * s00 - c0.z * s03 + c0.w * s04) * inv;
    o.c[1].z = -(c0.x * (c1.z * c3.w - c1.w * c3.z) - c0.z * (c1.x * c3.w - c1.w * c3.x) + c0.w * (c1.x * c3.z - c1.z * c3.x)) * inv;
    o.c[1].w = (c0.x * (c1.z * c2.w - c1.w * c2.z) - c0.z * (c1.x * c2.w - c1.w * c2.x) + c0.w * (c1.x * c2.z - c1.z * c2.x)) * inv;
    o.c[2].x = (c1.x * s01 - c1.y * s03 + c1.w * s05) * inv;
    o.c[2].y = -(c0.x * s01 - c0.y * s03 + c0.w * s05) * inv;
    o.c[2].z = (c0.x * (c1.y * c3.w - c1.w * c3.y) - c0.y * (c1.x * c3.w - c1.w * c3.x) + c0.w * (c1.x * c3.y - c1.y * c3.x)) * inv;
    o.c[2].w = -(c0.x * (c1.y * c2.w - c1.w * c2.y) - c0.y * (c1.x * c2.w - c1.w * c2.x) + c0.w * (c1.x * c2.y - c1.y * c2.x)) * inv;
    o.c[3].x = -(c1.x * s02 - c1.y * s04 + c1.z * s05) * inv;
    o.c[3].y = (c0.x * s02 - c0.y * s04 + c0.z * s05) * inv;
    o.c[3].z = -(c0.x * (c1.y * c3.z - c1.z * c3.y) - c0.y * (c1.x * c3.z - c1.z * c3.x) + c0.z * (c1.x * c3.y - c1.y * c3.x)) * inv;
    o.c[3].w = (c0.x * (c1.y * c2.z - c1.z * c2.y) - c0.y * (c1.x * c2.z - c1.z * c2.x) + c0.z * (c1.x * c2.y - c1.y * c2.x)) * inv;
    return o;
}
// glm::inverse for mat4 (vendored glm 1.0.0, detail/func_matrix.inl:347-405):
// cofactor pairs, sign pattern, determinant from the first row.
M4<float> GlmInverse(const M4<float>& mm) {
    auto m = [&](int c, int r) -> float { const V4<float>& v = mm.c[c]; return r == 0 ? v.x : r == 1 ? v.y : r == 2 ? v.z : v.w; };
    float C00 = m(2,2) * m(3,3) - m(3,2) * m(2,3), C02 = m(1,2) * m(3,3) - m(3,2) * m(1,3), C03 = m(1,2) * m(2,3) - m(2,2) * m(1,3);
    float C04 = m(2,1) * m(3,3) - m(3,1) * m(2,3), C06 = m(1,1) * m(3,3) - m(3,1) * m(1,3), C07 = m(1,1) * m(2,3) - m(2,1) * m(1,3);
    float C08 = m(2,1) * m(3,2) - m(3,1) * m(2,2), C10 = m(1,1) * m(3,2) - m(3,1) * m(1,2), C11 = m(1,1) * m(2,2) - m(2,1) * m(1,2);
    float C12 = m(2,0) * m(3,3) - m(3,0) * m(2,3), C14 = m(1,0) * m(3,3) - m(3,0) * m(1,3), C15 = m(1,0) * m(2,3) - m(2,0) * m(1,3);
    float C16 = m(2,0) * m(3,2) - m(3,0) * m(2,2), C18 = m(1,0) * m(3,2) - m(3,0) * m(1,2), C19 = m(1,0) * m(2,2) - m(2,0) * m(1,2);
    float C20 = m(2,0) * m(3,1) - m(3,0) * m(2,1), C22 = m(1,0) * m(3,1) - m(3,0) * m(1,1), C23 = m(1,0) * m(2,1) - m(2,0) * m(1,1);
    V4<float> F0(C00, C00, C02, C03), F1(C04, C04, C06, C07), F2(C08, C08, C10, C11);
    V4<float> F3(C12, C12, C14, C15), F4(C16, C16, C18, C19), F5(C20, C20, C22, C23);
    V4<float> A0(m(1,0), m(0,0), m(0,0), m(0,0)), A1(m(1,1), m(0,1), m(0,1), m(0,1));
    V4<float> A2(m(1,2), m(0,2), m(0,2), m(0,2)), A3(m(1,3), m(0,3), m(0,3), m(0,3));
    auto mul = [](V4<float> a, V4<float> b) { return V4<float>(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); };
    auto sub = [](V4<float> a, V4<float> b) { return V4<float>(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); };
    auto add = [](V4<float> a, V4<float> b) { return V4<float>(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); };
    V4<float> I0 = add(sub(mul(A1, F0), mul(A2, F1)), mul(A3, F2));
    V4<float> I1 = add(sub(mul(A0, F0), mul(A2, F3)), mul(A3, F4));
    V4<float> I2 = add(sub(mul(A0, F1), mul(A1, F3)), mul(A3, F5));
    V4<float> I3 = add(sub(mul(A0, F2), mul(A1, F4)), mul(A2, F5));
    V4<float> SA(1, -1, 1, -1), SB(-1, 1, -1, 1);
    M4<float> inv;
    inv.c[0] = mul(I0, SA); inv.c[1] = mul(I1, SB); inv.c[2] = mul(I2, SA); inv.c[3] = mul(I3, SB);
    V4<float> row0(inv.c[0].x, inv.c[1].x, inv.c[2].x, inv.c[3].x);
    V4<float> d0 = mul(mm.c[0], row0);
    float d1 = (d0.x + d0.y) + (d0.z + d0.w);
    float one_over = 1.0f / d1;
    for (int i = 0; i < 4; ++i) inv.c[i] = V4<float>(inv.c[i].x * one_over, inv.c[i].y * one_over, inv.c[i].z * one_over, inv.c[i].w * one_over);
    return inv;
}
M4<float> LoadM4(const float* p) {
    M4<float> m;
    for (int i = 0; i < 4; ++i) m.c[i] = V4<float>(p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]);
    return m;
}
void StoreM4(const M4<float>& m, float* p) {
    for (int i = 0; i < 4; ++i) { p[4 * i] = m.c[i].x; p[4 * i + 1] = m.c[i].y; p[4 * i + 2] = m.c[i].z; p[4 * i + 3] = m.c[i].w; }
}
template <class R> M4<R> CvtM4(const M4<float>& m) {
    M4<R> o;
    for (int i = 0; i < 4; ++i) o.c[i] = V4<R>(R(m.c[i].x), R(m.c[i].y), R(m.c[i].z), R(m.c[i].w));
    return o;
}

template <class R> struct M3 {
    V3<R> c[3];
    M3() { c[0] = V3<R>(1, 0, 0); c[1] = V3<R>(0, 1, 0); c[2] = V3<R>(0, 0, 1); }
};
template <class R> V3<R> Mul3(const M3<R>& m, V3<R> v) {  // matrix_maths.cuh:91-99
    return V3<R>(m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z,
                 m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z,
                 m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z);
}
template <class R> M3<R> Transpose3(const M3<R>& m) {
    M3<R> t;
    t.c[0] = V3<R>(m.c[0].x, m.c[1].x, m.c[2].x);
    t.c[1] = V3<R>(m.c[0].y, m.c[1].y, m.c[2].y);
    t.c[2] = V3<R>(m.c[0].z, m.c[1].z, m.c[2].z);
    return t;
}
// generateONBFrisvad, matrix_maths.cuh:375-396
template <class R> M3<R> OnbFrisvad(V3<R> n) {
    M3<R> ret;
    ret.c[1] = n;
    if (n.z < R(-0.999805696f)) {
        ret.c[0] = V3<R>(0, -1, 0);
        ret.c[2] = V3<R>(-1, 0, 0);
    } else {
        R a = R(1) / (R(1) + n.z);
        R b = -n.x * n.y * a;
        ret.c[0] = V3<R>(R(1) - n.x * n.x * a, b, -n.x);
        ret.c[2] = V3<R>(b, R(1) - n.y * n.y * a, -n.y);
    }
    ret.c[1] = ret.c[2];
    ret.c[2] = n;
    return ret;
}

// ------------------------------------------------------------------ constants
// constants.cuh:9-24, atmosphere.cuh:67-69, 292-310
const float PI_F = 3.14159265358979323846f;
const float INV_PI_F = 0.31830988618379067154f;
const float INV_UINT32MAX_F = 1.0f / float(UINT32_MAX);
const float INV_255_F = 0.00392156862f;
const float TRI_EPS_F = 1.0e-5f;
const float HIT_EPS_F = 0.001f;
const float SOLAR_DIST_F = 149.597e9f;
const float SUN_NITS_F = 6.0e5f;
// 6'378.1370_km and 6'438.1370_km through operator""_km (constants.cuh:27-29):
// both products are integers that float represents exactly.
const float EARTH_R_F = 6378137.0f;
const float ATMOS_R_F = 6438137.0f;
const float HR_F = 7994.0f;
const float HM_F = 1200.0f;

// ------------------------------------------------------------------ scene
struct Bounds3f {  // bounds.cuh:5-60
    V3<float> pmin{FLT_MAX, FLT_MAX, FLT_MAX}, pmax{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float surfaceArea() const {
        V3<float> d = pmax - pmin;
        return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
    void grow(V3<float> p) { pmin = Fmin3(pmin, p); pmax = Fmax3(pmax, p); }
    void grow(const Bounds3f& b) { grow(b.pmax); grow(b.pmin); }
};

struct Triangle {  // triangle.cuh:54-129
    V3<float> p[3];
    V3<float> n[3];
    V2<float> t[3];
    V3<float> local_geometric_normal;
    int32_t material_id = -1, light_id = -1, mesh_id = -1;
};

struct BVH2Node {  // blas.cuh:9-35 (32 B)
    int32_t first = -1;   // left child or first tri-index slot
    uint32_t count = 0;
    Bounds3f bounds;
    bool isLeaf() const { return count > 0; }
    float surfaceArea() const { return count <= 0 ? 0.0f : bounds.surfaceArea(); }
};
static_assert(sizeof(BVH2Node) == 32, "BVH2Node layout");

struct TLASNode {  // blas.cuh:209-238 (32 B)
    Bounds3f bounds;
    uint32_t left_right_index = 0;
    int32_t blas_index = -1;
};
static_assert(sizeof(TLASNode) == 32, "TLASNode layout");

struct BLAS2 {  // blas.cuh:45-207
    Bounds3f bounds, original_bounds;
    int64_t root = -1;
    int32_t mesh_id = -1;
    M4<float> inv_model;
};

struct TriangleMesh {  // triangle.cuh:132-154
    M4<float> curr_inv_model, prev_inv_model;
    int32_t blas_index = -1;
    uint32_t prim_count = 0;
    int32_t prim_offset = -1;
};

struct Material {  // material.cuh:21-75
    int albedo_texture_id = -1;
    V3<float> albedo{0.8f, 0.8f, 0.8f};
    int orm_texture_id = -1;
    float metallic = 0.0f, roughness = 0.5f;
    int transmission_texture_id = -1;
    float transmission = 0.0f;
    float ior = 1.45f;
    int emission_texture_id = -1;
    V3<float> emissive_factor{0, 0, 0};
    float emission_scale_nits = 1.0f;
    int normal_texture_id = -1;
    float normal_scale = 1.0f;
};

struct Texture {  // texture.cuh:21-65
    int width = 0, height = 0, channel_count = 3, bit_depth = 8;
    int64_t pixel_buffer_index = -1;
};

struct Light {  // light.cuh:68-95 (+ the instance the triangle belongs to)
    float area = 0.0f;
    int32_t prim_id = -1;
    int32_t inst_id = -1;
};

}  // namespace orc

struct orc_scene {
    std::vector<orc::Triangle> tris;
    std::vector<int32_t> tri_index;
    std::vector<orc::BVH2Node> nodes;
    std::vector<orc::TLASNode> tlas_nodes;
    orc::Bounds3f tlas_bounds;
    int32_t tlas_root = -1;
    std::vector<orc::BLAS2> blas;               // one per instance
    std::vector<orc::TriangleMesh> meshes;      // one per instance (shading matrices)
    std::vector<int64_t> geom_root;             // BLAS root per geometry mesh
    std::vector<int32_t> geom_first, geom_count;
    std::vector<int32_t> inst_geom;
    std::vector<orc::Material> materials;
    std::vector<orc::Texture> textures;
    std::vector<uint8_t> pixels;
    std::vector<orc::Light> lights;
    int32_t max_blas_depth = 0, max_tlas_depth = 0;
};

namespace orc {

// ------------------------------------------------------------------ builders
// BLASBuilder, as_builder.cuh:11-295 (centroid cache, 8-bin SAH over 3 axes,
// leaves of <= 8 triangles, children allocated as a (left, left+1) pair).
struct BlasBuilder {
    orc_scene* S;
    std::vector<V3<float>> cache;   // indexed by global triangle id
    float comp(V3<float> v, int idx) { return idx == 0 ? v.x : idx == 1 ? v.y : v.z; }

    void updateNodeBounds(uint32_t ni) {  // :49-62
        BVH2Node& node = S->nodes[ni];
        node.bounds.pmin = V3<float>(FLT_MAX, FLT_MAX, FLT_MAX);
        node.bounds.pmax = V3<float>(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        for (uint32_t first = node.first, i = 0; i < node.count; i++) {
            const Triangle& t = S->tris[S->tri_index[first + i]];
            node.bounds.grow(t.p[0]); node.bounds.grow(t.p[1]); node.bounds.grow(t.p[2]);
        }
    }
    struct Bin { Bounds3f bounds; uint32_t count = 0; };
    float findBestSplitPlane(const BVH2Node& node, int* axis, float* split_pos) {  // :110-179
        const int BINS = 8;
        float best_cost = INFINITY;
        for (int a = 0; a < 3; a++) {
            float bmin = INFINITY, bmax = -INFINITY;
            for (uint32_t i = 0; i < node.count; i++) {
                V3<float> c = cache[S->tri_index[node.first + i]];
                bmin = std::fmin(bmin, comp(c, a));
                bmax = std::fmax(bmax, comp(c, a));
            }
            if (bmin == bmax) continue;
            Bin bin[BINS];
            float scale = float(BINS) / (bmax - bmin);
            for (uint32_t i = 0; i < node.count; i++) {
                int32_t tid = S->tri_index[node.first + i];
                V3<float> c = cache[tid];
                const Triangle& t = S->tris[tid];
                int bi = std::min(BINS - 1, (int32_t)((comp(c, a) - bmin) * scale));
                bin[bi].count++;
                bin[bi].bounds.grow(t.p[0]); bin[bi].bounds.grow(t.p[1]); bin[bi].bounds.grow(t.p[2]);
            }
            const int P = BINS - 1;
            float la[P], ra[P];
            int32_t lc[P], rc[P];
            Bounds3f lb, rb;
            int32_t ls = 0, rs = 0;
            for (int i = 0; i < P; i++) {
                ls += bin[i].count; lc[i] = ls; lb.grow(bin[i].bounds); la[i] = lb.surfaceArea();
                rs += bin[P - i].count; rc[P - 1 - i] = rs; rb.grow(bin[P - i].bounds); ra[P - 1 - i] = rb.surfaceArea();
            }
            scale = (bmax - bmin) / float(BINS);
            for (int i = 0; i < P; i++) {
                float cost = lc[i] * la[i] + rc[i] * ra[i];
                if (cost < best_cost) { *axis = a; *split_pos = bmin + scale * (i + 1); best_cost = cost; }
            }
        }
        return best_cost;
    }
    void subdivide(uint32_t nid, uint32_t* next, int depth, int* max_depth) {  // :181-233
        if (depth > *max_depth) *max_depth = depth;
        BVH2Node& parent = S->nodes[nid];
        int axis = -1;
        float split_pos = 0;
        findBestSplitPlane(parent, &axis, &split_pos);
        if (parent.count <= 8) return;
        int32_t i = parent.first;
        int32_t j = i + parent.count - 1;
        while (i <= j) {
            if (comp(cache[S->tri_index[i]], axis) < split_pos) i++;
            else std::swap(S->tri_index[i], S->tri_index[j--]);
        }
        int32_t split = i;
        BVH2Node left, right;
        left.count = split - parent.first;
        if (left.count == 0 || left.count == parent.count) return;
        int32_t lid = (*next)++;
        int32_t rid = (*next)++;
        left.first = parent.first;
        right.first = split;
        right.count = parent.count - left.count;
        parent.count = 0;
        parent.first = lid;
        S->nodes[lid] = left;
        S->nodes[rid] = right;
        updateNodeBounds(lid);
        updateNodeBounds(rid);
        subdivide(lid, next, depth + 1, max_depth);
        subdivide(rid, next, depth + 1, max_depth);
    }
    // buildBLAS, :235-261.  The node range reserved is 2N-1 as in the reference.
    int64_t build(int32_t first, int32_t count, Bounds3f* bounds, int* max_depth) {
        for (int32_t i = 0; i < count; i++) {
            const Triangle& t = S->tris[first + i];
            cache[first + i] = (t.p[0] + t.p[1] + t.p[2]) * 0.3333f;
        }
        int32_t root = (int32_t)S->nodes.size();
        S->nodes.resize(S->nodes.size() + (2 * (size_t)count - 1));
        size_t ti0 = S->tri_index.size();
        S->tri_index.resize(ti0 + count);
        std::iota(S->tri_index.begin() + ti0, S->tri_index.end(), first);
        BVH2Node& r = S->nodes[root];
        r.count = count;
        r.first = first;
        updateNodeBounds(root);
        *bounds = S->nodes[root].bounds;
        uint32_t next = root + 1;
        subdivide(root, &next, 0, max_depth);
        return root;
    }
};

// BLAS2::setTransform, blas.cuh:190-200
void BlasSetTransform(BLAS2& b, const M4<float>& model) {
    b.inv_model = Inverse(model);
    V3<float> bmin = b.original_bounds.pmin, bmax = b.original_bounds.pmax;
    b.bounds = Bounds3f();
    for (int i = 0; i < 8; i++) {
        V4<float> c(i & 1 ? bmax.x : bmin.x, i & 2 ? bmax.y : bmin.y, i & 4 ? bmax.z : bmin.z, 1.0f);
        b.bounds.grow(xyz(Mul(model, c)));
    }
}

// TLASBuilder, as_builder.cuh:297-392: agglomerative clustering (findBestMatch
// over half surface areas), interior children packed 16|16 bits, root copied
// to node 0.
int32_t FindBestMatch(const orc_scene* S, const std::vector<TLASNode>& nodes, const int32_t* ids, int32_t n, int32_t A) {
    float smallest = INFINITY;
    int32_t best = -1;
    for (int32_t B = 0; B < n; B++) {
        if (B == A) continue;
        V3<float> bmax = Fmax3(nodes[ids[A]].bounds.pmax, nodes[ids[B]].bounds.pmax);
        V3<float> bmin = Fmin3(nodes[ids[A]].bounds.pmin, nodes[ids[B]].bounds.pmin);
        V3<float> e = bmax - bmin;
        float sa = e.x * e.y + e.y * e.z + e.z * e.x;
        if (sa < smallest) { smallest = sa; best = B; }
    }
    (void)S;
    return best;
}
void BuildTlas(orc_scene* S) {
    const uint32_t N = (uint32_t)S->blas.size();
    S->tlas_nodes.clear();
    if (N == 0) { S->tlas_root = -1; S->tlas_bounds = Bounds3f(); return; }
    std::vector<TLASNode>& nodes = S->tlas_nodes;
    nodes.resize(2 * N);
    std::vector<int32_t> ids(N);
    int32_t ptr = 1;
    int32_t left = N;
    for (uint32_t i = 0; i < N; i++) {
        ids[i] = ptr;
        nodes[ptr].bounds = S->blas[i].bounds;
        nodes[ptr].blas_index = i;
        nodes[ptr++].left_right_index = 0u;
    }
    int A = 0, B = FindBestMatch(S, nodes, ids.data(), left, A);
    while (left > 1) {
        int C = FindBestMatch(S, nodes, ids.data(), left, B);
        if (A == C) {
            int32_t na = ids[A], nb = ids[B];
            TLASNode nn;
            nn.left_right_index = (uint32_t)na + ((uint32_t)nb << 16);
            nn.bounds.pmin = Fmin3(nodes[na].bounds.pmin, nodes[nb].bounds.pmin);
            nn.bounds.pmax = Fmax3(nodes[na].bounds.pmax, nodes[nb].bounds.pmax);
            nodes[ptr] = nn;
            ids[A] = ptr++;
            ids[B] = ids[left - 1];
            B = FindBestMatch(S, nodes, ids.data(), --left, A);
        } else {
            A = B; B = C;
        }
    }
    nodes[0] = nodes[ids[A]];
    S->tlas_root = 0;
    S->tlas_bounds = nodes[0].bounds;
    nodes.resize(ptr);
    // depth statistic
    int maxd = 0;
    std::vector<std::pair<uint32_t, int>> st{{0u, 0}};
    while (!st.empty()) {
        auto [n, d] = st.back(); st.pop_back();
        maxd = std::max(maxd, d);
        if (nodes[n].left_right_index != 0) {
            st.push_back({nodes[n].left_right_index & 0xFFFF, d + 1});
            st.push_back({(nodes[n].left_right_index >> 16) & 0xFFFF, d + 1});
        }
    }
    S->max_tlas_depth = maxd;
}

// ------------------------------------------------------------------ rays
template <class R> struct Ray {  // ray.cuh:6-34
    V3<R> o, d, inv;
    Ray() = default;
    Ray(V3<R> org, V3<R> dir) : o(org), d(dir), inv(R(1) / dir) {}
    V3<R> at(R t) const { return o + (t * d); }
};
template <class R> Ray<R> Transform(const Ray<R>& r, const M4<R>& m) {
    return Ray<R>(xyz(Mul(m, V4<R>(r.o, R(1)))), xyz(Mul(m, V4<R>(r.d, R(0)))));
}
// Bounds3f::intersectP, bounds.cuh:38-57
template <class R> bool SlabHit(const Bounds3f& b, const Ray<R>& ray, R tmin, R tmax, R* h0, R* h1) {
    V3<R> t0 = (cvt<R>(b.pmin) - ray.o) * ray.inv;
    V3<R> t1 = (cvt<R>(b.pmax) - ray.o) * ray.inv;
    V3<R> lo = Fmin3(t0, t1);
    V3<R> hi = Fmax3(t0, t1);
    R tenter = Fmax(Fmax(lo.x, lo.y), lo.z);
    R texit = Fmin(Fmin(hi.x, hi.y), hi.z);
    bool hit = (tenter <= texit) && (texit >= tmin) && (tenter <= tmax);
    if (hit) { *h0 = tenter; *h1 = texit; }
    return hit;
}
template <class R> struct Isect {  // interaction.cuh:69-88
    R distance = Inf<R>();
    int32_t prim = -1, inst = -1;
    V3<R> bary;
};
// Triangle::intersect, triangle.cuh:66-102 (Moller-Trumbore)
template <class R> void TriIntersect(const Triangle& tri, const Ray<R>& ray, R tmin, R tmax, Isect<R>* it) {
    it->distance = Inf<R>();
    V3<R> v0 = cvt<R>(tri.p[0]);
    V3<R> e1 = cvt<R>(tri.p[1]) - v0;
    V3<R> e2 = cvt<R>(tri.p[2]) - v0;
    V3<R> pvec = Cross(ray.d, e2);
    R det = Dot(e1, pvec);
    if (det > -R(TRI_EPS_F) && det < R(TRI_EPS_F)) return;
    R invDet = R(1) / det;
    V3<R> tvec = ray.o - v0;
    R u = invDet * Dot(tvec, pvec);
    if (u < R(0) || u > R(1)) return;
    V3<R> qvec = Cross(tvec, e1);
    R v = invDet * Dot(ray.d, qvec);
    if (v < R(0) || u + v > R(1)) return;
    R t = invDet * Dot(e2, qvec);
    if (t > R(TRI_EPS_F) && t < tmax && t > tmin) {
        it->distance = t;
        it->bary = V3<R>(R(1) - u - v, u, v);
    }
}

// max_stack: the deepest the TLAS and BLAS stacks got together (tlas_depth
// holds the TLAS stack's size while a BLAS is traversed).
struct Counters {
    uint64_t interior = 0, leaf = 0, tri = 0, entries = 0, tlas = 0;
    int max_stack = 0, tlas_depth = 0;
    void depth(int d) { if (d > max_stack) max_stack = d; }
};
// Traversal events of every trace a path makes while g_events is set
// (orc_sample_events): closest-hit and shadow rays alike, and the ray count.
struct SampleEvents {
    uint64_t interior = 0, leaf = 0, tri = 0, entries = 0, tlas = 0, rays = 0;
};
thread_local SampleEvents* g_events = nullptr;
inline void AddEvents(const Counters& c) {
    if (!g_events) return;
    g_events->interior += c.interior;
    g_events->leaf += c.leaf;
    g_events->tri += c.tri;
    g_events->entries += c.entries;
    g_events->tlas += c.tlas;
    g_events->rays += 1;
}

// BLAS2::intersect, blas.cuh:113-188 (unbounded stack: the reference's 16
// entries overflow on deep trees, which is undefined behaviour there).
// init: the distance the BLAS's own closest hit starts from (+inf in the
// reference; the closest hit so far in the GPU kernels' bound, see
// TlasIntersect).
template <class R> Isect<R> BlasIntersect(const orc_scene* S, const BLAS2& b, const Ray<R>& wray, R tmin, R tmax, Counters* cn,
                                          R init = Inf<R>()) {
    Isect<R> closest;
    closest.distance = init;
    if (b.root < 0) return closest;
    Ray<R> ray = Transform(wray, CvtM4<R>(b.inv_model));
    std::vector<uint32_t> stack;
    stack.reserve(64);
    stack.push_back((uint32_t)b.root);
    Isect<R> it;
    while (!stack.empty()) {
        cn->depth(cn->tlas_depth + (int)stack.size());
        const BVH2Node& node = S->nodes[stack.back()];
        stack.pop_back();
        if (!node.isLeaf()) {
            cn->interior++;
            R c1 = Inf<R>(), c2 = Inf<R>(), x1 = 0, x2 = 0;
            bool h1 = SlabHit(S->nodes[node.first].bounds, ray, tmin, tmax, &c1, &x1);
            bool h2 = SlabHit(S->nodes[node.first + 1].bounds, ray, tmin, tmax, &c2, &x2);
            if (c1 > c2) {
                if (h1 && c1 < closest.distance) stack.push_back(node.first);
                if (h2 && c2 < closest.distance) stack.push_back(node.first + 1);
            } else {
                if (h2 && c2 < closest.distance) stack.push_back(node.first + 1);
                if (h1 && c1 < closest.distance) stack.push_back(node.first);
            }
        } else {
            cn->leaf++;
            for (uint32_t i = 0; i < node.count; i++) {
                int32_t prim = S->tri_index[node.first + i];
                cn->tri++;
                TriIntersect(S->tris[prim], ray, tmin, tmax, &it);
                if (it.distance < Inf<R>() && it.distance < tmax && it.distance > tmin && it.distance < closest.distance) {
                    closest = it;
                    closest.prim = prim;
                    closest.inst = b.mesh_id;
                    tmax = closest.distance;
                }
            }
        }
    }
    return closest;
}
// BLAS2::intersectP, blas.cuh:50-111
template <class R> bool BlasIntersectP(const orc_scene* S, const BLAS2& b, const Ray<R>& wray, R tmin, R tmax, Counters* cn) {
    if (b.root < 0) return false;
    Ray<R> ray = Transform(wray, CvtM4<R>(b.inv_model));
    std::vector<uint32_t> stack;
    stack.reserve(64);
    stack.push_back((uint32_t)b.root);
    Isect<R> it;
    while (!stack.empty()) {
        cn->depth(cn->tlas_depth + (int)stack.size());
        const BVH2Node& node = S->nodes[stack.back()];
        stack.pop_back();
        if (!node.isLeaf()) {
            cn->interior++;
            R c1 = Inf<R>(), c2 = Inf<R>(), x1 = 0, x2 = 0;
            bool h1 = SlabHit(S->nodes[node.first].bounds, ray, tmin, tmax, &c1, &x1);
            bool h2 = SlabHit(S->nodes[node.first + 1].bounds, ray, tmin, tmax, &c2, &x2);
            if (h1) stack.push_back(node.first);
            if (h2) stack.push_back(node.first + 1);
        } else {
            cn->leaf++;
            for (uint32_t i = 0; i < node.count; i++) {
                int32_t prim = S->tri_index[node.first + i];
                cn->tri++;
                TriIntersect(S->tris[prim], ray, tmin, tmax, &it);
                if (it.distance < Inf<R>() && it.distance < tmax && it.distance > tmin) return true;
            }
        }
    }
    return false;
}
// TLAS::intersect, tlas.cu:60-125.  entry_bound: enter each BLAS with
// tmax = min(tmax, closest so far) and cull its nodes against that closest
// distance, as the GPU kernels do (not the reference; a test compares the
// two).
template <class R> Isect<R> TlasIntersect(const orc_scene* S, const Ray<R>& ray, R tmin, R tmax, Counters* cn,
                                          bool entry_bound = false) {
    PhaseScope<R> phase_(PH_TRAV);
    if (S->tlas_root < 0) return Isect<R>();
    R c1 = Inf<R>(), c2 = Inf<R>();
    if (!SlabHit(S->tlas_bounds, ray, tmin, tmax, &c1, &c2)) return Isect<R>();
    c1 = Inf<R>(); c2 = Inf<R>();
    std::vector<int32_t> stack;
    stack.reserve(32);
    stack.push_back(S->tlas_root);
    Isect<R> closest;
    while (!stack.empty()) {
        cn->depth((int)stack.size());
        const TLASNode& node = S->tlas_nodes[stack.back()];
        stack.pop_back();
        cn->tlas_depth = (int)stack.size();
        cn->tlas++;
        if (node.left_right_index != 0) {
            uint32_t i1 = node.left_right_index & 0xFFFF, i2 = (node.left_right_index >> 16) & 0xFFFF;
            R x1 = 0, x2 = 0;
            bool h1 = SlabHit(S->tlas_nodes[i1].bounds, ray, tmin, tmax, &c1, &x1);
            bool h2 = SlabHit(S->tlas_nodes[i2].bounds, ray, tmin, tmax, &c2, &x2);
            if (c1 > c2) {
                if (h1 && c1 < closest.distance) stack.push_back(i1);
                if (h2 && c2 < closest.distance) stack.push_back(i2);
            } else {
                if (h2 && c2 < closest.distance) stack.push_back(i2);
                if (h1 && c1 < closest.distance) stack.push_back(i1);
            }
        } else {
            cn->entries++;
            Isect<R> it = entry_bound
                              ? BlasIntersect(S, S->blas[node.blas_index], ray, tmin, Fmin(tmax, closest.distance), cn,
                                              closest.distance)
                              : BlasIntersect(S, S->blas[node.blas_index], ray, tmin, tmax, cn);
            if (it.distance < closest.distance) closest = it;
        }
    }
    return closest;
}
// TLAS::intersectP, tlas.cu:8-58
template <class R> bool TlasIntersectP(const orc_scene* S, const Ray<R>& ray, R tmin, R tmax, Counters* cn) {
    PhaseScope<R> phase_(PH_TRAV);
    if (S->tlas_root < 0) return false;
    R c1 = Inf<R>(), c2 = Inf<R>();
    if (!SlabHit(S->tlas_bounds, ray, tmin, tmax, &c1, &c2)) return false;
    std::vector<int32_t> stack;
    stack.reserve(32);
    stack.push_back(S->tlas_root);
    while (!stack.empty()) {
        cn->depth((int)stack.size());
        const TLASNode& node = S->tlas_nodes[stack.back()];
        stack.pop_back();
        cn->tlas_depth = (int)stack.size();
        cn->tlas++;
        if (node.left_right_index != 0) {
            uint32_t i1 = node.left_right_index & 0xFFFF, i2 = (node.left_right_index >> 16) & 0xFFFF;
            R x1 = 0, x2 = 0;
            bool h1 = SlabHit(S->tlas_nodes[i1].bounds, ray, tmin, tmax, &c1, &x1);
            bool h2 = SlabHit(S->tlas_nodes[i2].bounds, ray, tmin, tmax, &c2, &x2);
            if (h1) stack.push_back(i1);
            if (h2) stack.push_back(i2);
        } else {
            cn->entries++;
            if (BlasIntersectP(S, S->blas[node.blas_index], ray, tmin, tmax, cn)) return true;
        }
    }
    return false;
}
// INTERSECT_DEBUG brute force, integrator.cuh:64-122
template <class R> Isect<R> BruteClosest(const orc_scene* S, const Ray<R>& ray, R tmin, R tmax) {
    Isect<R> closest, it;
    for (size_t inst = 0; inst < S->meshes.size(); inst++) {
        const TriangleMesh& m = S->meshes[inst];
        Ray<R> oray = Transform(ray, CvtM4<R>(S->blas[inst].inv_model));
        for (int32_t p = m.prim_offset; p < m.prim_offset + (int32_t)m.prim_count; p++) {
            TriIntersect(S->tris[p], oray, tmin, tmax, &it);
            if (it.distance < Inf<R>() && it.distance < tmax && it.distance > tmin && it.distance < closest.distance) {
                closest = it; closest.prim = p; closest.inst = (int32_t)inst;
            }
        }
    }
    return closest;
}
template <class R> bool BruteAny(const orc_scene* S, const Ray<R>& ray, R tmin, R tmax) {
    Isect<R> it;
    for (size_t inst = 0; inst < S->meshes.size(); inst++) {
        const TriangleMesh& m = S->meshes[inst];
        Ray<R> oray = Transform(ray, CvtM4<R>(S->blas[inst].inv_model));
        for (int32_t p = m.prim_offset; p < m.prim_offset + (int32_t)m.prim_count; p++) {
            TriIntersect(S->tris[p], oray, tmin, tmax, &it);
            if (it.distance < Inf<R>() && it.distance < tmax && it.distance > tmin) return true;
        }
    }
    return false;
}

// ------------------------------------------------------------------ sampling
// RNG::pcg_hash + IndependentSampler, samplers.cuh:10-48
inline uint32_t PcgHash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
template <class R> struct Sampler {
    uint32_t seed = 0;
    void initPixelSeed(int px, int py, int width, int index, int mode) {
        if (mode == KPT_SEED_XOR) {
            seed = (uint32_t)(px + py * width) ^ (uint32_t)index;
            return;
        }
        seed = (uint32_t)(px + py * width);
        seed = (uint32_t)((float)seed * ((index == 0) ? 1.0f : (float)index));
        seed += 0;
    }
    R get1D() { seed = PcgHash(seed); return R((float)seed * INV_UINT32MAX_F); }
    V2<R> get2D() { V2<R> v; v.x = get1D(); v.y = get1D(); return v; }
};

// ------------------------------------------------------------------ BSDF
template <class R> using Mt = Mth<R>;
template <class R> R ComputeF0(R ior) { return Sqr((ior - R(1)) / (ior + R(1))); }

template <class R> struct GGX {  // bsdf.cuh:18-88
    R alpha = R(0.25f);
    bool smooth() const { return alpha < R(0.001f); }
    R D(V3<R> wm) const {
        R a = alpha;
        if (smooth()) a = Fmax(a, R(1.0e-3f));
        R a2 = Sqr(a);
        R ndh = Sqr(Saturate(wm.z));
        R b = R((double)ndh * ((double)a2 - 1.0) + 1.0);
        return a2 / (R(PI_F) * Sqr(b));
    }
    R G(V3<R> wo, V3<R> wi) const {
        R nv = Fmax(wo.z, R(0));
        R nl = Fabs(wi.z);
        R tv = Sqrt(Sqr(alpha) + (R(1) - Sqr(alpha)) * Sqr(nv));
        R tl = Sqrt(Sqr(alpha) + (R(1) - Sqr(alpha)) * Sqr(nl));
        R g1v = (R(2) * nv) / (nv + tv);
        R g1l = (R(2) * nl) / (nl + tl);
        return g1v * g1l;
    }
    V3<R> sampleWm(V3<R> w, V2<R> u) const {
        (void)w;
        R a2 = Sqr(alpha);
        R e0 = u.x, e1 = u.y;
        R theta = Mt<R>::acos(Sqrt((R(1) - e0) / ((a2 - R(1)) * e0 + R(1))));
        R phi = R(2) * R(PI_F) * e1;
        // sphericalToCartesian, vector_maths.cuh:37-44
        V3<R> wm(Mt<R>::sin(theta) * Mt<R>::cos(phi), Mt<R>::sin(theta) * Mt<R>::sin(phi), Mt<R>::cos(theta));
        return Normalize(wm);
    }
    void regularize() { if (alpha < R(0.3f)) alpha = Clamp(R(2) * alpha, R(0.1f), R(0.3f)); }
};

enum Scatter { Absorbed = 0, Emitted = 1, Reflected = 2, Transmitted = 4, Diffuse = 8, Glossy = 16, Specular = 32 };
template <class R> struct BSDFSample { int scatter = Absorbed; V3<R> f; V3<R> wi; R pdf = 0; };

// fresnelDielectric, bsdf.cuh:122-141
template <class R> R FresnelDielectric(R cosT, R ior) {
    cosT = Clamp(cosT, R(-1), R(1));
    if (cosT < R(0)) { ior = R(1) / ior; cosT = -cosT; }
    R s2 = (R(1) - cosT * cosT);
    R s2t = s2 / (ior * ior);
    if (s2t >= R(1)) return R(1);
    R ct = Sqrt(R(1) - s2t);
    R rp = (ior * cosT - ct) / (ior * cosT + ct);
    R rs = (cosT - ior * ct) / (cosT + ior * ct);
    return (rp * rp + rs * rs) * R(0.5f);
}
// fresnelSchlick, bsdf.cuh:144-148
template <class R> V3<R> FresnelSchlick(R c, V3<R> F0) {
    V3<R> F = F0 + (R(1) - F0) * Pow5(R(1) - c);
    return Clamp3(F, V3<R>(0, 0, 0), V3<R>(1, 1, 1));
}
// microfacetBRDF, bsdf.cuh:151-164
template <class R> R MicrofacetBRDF(V3<R> wo, V3<R> wi, V3<R> wm, const GGX<R>& g) {
    R nv = Saturate(wo.z);
    R nl = Saturate(wi.z);
    R D = g.D(wm);
    R G = g.G(wo, wi);
    R f = (D * G) / (R(4) * Fmax(nv, R(1.0e-3f)) * nl);
    if (wi.z <= R(0) || Dot(wi, wm) <= R(0)) f = 0;
    return f;
}
// refract, vector_maths.cuh:46-65
template <class R> bool Refract(V3<R> wi, V3<R> n, R ior, V3<R>* wt) {
    R c = Dot(wi, n);
    if (c < R(0)) { ior = R(1) / ior; c *= R(-1); n = n * R(-1); }
    R s2 = (R(1) - c * c);
    R s2t = s2 / (ior * ior);
    if (s2t >= R(1)) return false;
    R ct = Sqrt(R(1) - s2t);
    *wt = (R(-1) * wi) / ior + (c / ior - ct) * n;
    return true;
}
template <class R> bool SameHemi(V3<R> a, V3<R> b, V3<R> n) { return (Dot(a, n) * Dot(b, n)) > R(0); }
template <class R> bool SameHemi(V3<R> a, V3<R> b) { return Dot(a, b) > R(0); }

// UnifiedBSDF, bsdf.cuh:166-624
template <class R> struct BSDF {
    GGX<R> ggx;
    M3<R> basis;
    V3<R> albedo{1, 0, 1};
    R roughness = R(0.5f), metallic = 0, transmission = 0, ior = R(1.45f);
    bool backface = false;

    R fDiffuse(V3<R> wo, V3<R> wi) const { R s = (wo.z * wi.z > R(0)) ? R(1) : R(0); return s * R(INV_PI_F); }
    R pdfDiffuse(V3<R> wo, V3<R> wi) const { (void)wo; return wi.z * R(INV_PI_F); }
    V3<R> fGlossy(V3<R> wo, V3<R> wi, V3<R> wm) const {
        R M = MicrofacetBRDF(wo, wi, wm, ggx);
        R F = FresnelDielectric(Dot(wo, wm), ior);
        R v = F * M;
        return V3<R>(v, v, v);
    }
    R pdfGlossy(V3<R> wo, V3<R> wi, V3<R> wm) const {
        (void)wi;
        R voh = Fmax(Dot(wo, wm), R(0));
        R noh = Fmax(wm.z, R(0));
        R D = ggx.D(wm);
        return (D * noh) / (R(4) * voh);
    }
    R pdfBTDF(V3<R> wo, V3<R> wi, V3<R> ht, R eta) const {
        R temp = Dot(wi, ht) * eta + Dot(wo, ht);
        R dw = AbsDot(wi, ht) / (temp * temp);
        R noh = Fabs(ht.z);
        R D = ggx.D(ht);
        R F = FresnelDielectric(AbsDot(wo, ht), eta);
        return (D * noh) * dw * (R(1) - F);
    }
    V3<R> fBTDF(V3<R> wo, V3<R> wi, V3<R> ht, R eta) const {
        R temp = Dot(wi, ht) * eta + Dot(wo, ht);
        R F = FresnelDielectric(AbsDot(wo, ht), eta);
        R T = ggx.D(ht) * ggx.G(wo, wi) * Fabs(Dot(wi, ht) * Dot(wo, ht) / (wi.z * wo.z * temp * temp));
        return albedo * (R(1) - F) * T;
    }
    V3<R> fOpaque(V3<R> wo, V3<R> wi) const {
        V3<R> wm = Normalize(wo + wi);
        V3<R> g = fGlossy(wo, wi, wm);
        V3<R> wd = R(1) - g;
        V3<R> diff = wd * (albedo * fDiffuse(wo, wi));
        return diff + g;
    }
    R pdfOpaque(V3<R> wo, V3<R> wi) const {
        V3<R> wm = Normalize(wo + wi);
        R ps = pdfGlossy(wo, wi, wm);
        R pd = pdfDiffuse(wo, wi);
        return R(0.5f) * (ps + pd);
    }
    V3<R> fConductor(V3<R> wo, V3<R> wi) const {
        V3<R> wm = Normalize(wo + wi);
        R nov = Clamp(wo.z, R(0), R(1));
        R nol = Clamp(wi.z, R(0), R(1));
        if (nov == R(0) || nol == R(0)) return V3<R>(0, 0, 0);
        R voh = Clamp(Dot(wo, wm), R(0), R(1));
        V3<R> M(0, 0, 0);
        if (nol > R(0) && voh > R(0)) {
            V3<R> F = FresnelSchlick(voh, albedo);
            M = F * MicrofacetBRDF(wo, wi, wm, ggx);
        }
        return M;
    }
    R pdfConductor(V3<R> wo, V3<R> wi) const { V3<R> wm = Normalize(wo + wi); return pdfGlossy(wo, wi, wm); }
    V3<R> fTransparent(V3<R> wo, V3<R> wi) const {
        R co = wo.z, ci = wi.z;
        bool refl = (co * ci) > R(0);
        R eta = R(1);
        if (!refl) eta = (!backface) ? ior : (R(1) / ior);
        V3<R> wm = (eta * wi + wo);
        wm = Normalize((wm.z > R(0)) ? wm : -wm);
        if ((Dot(wm, wi) * ci < R(0)) || (Dot(wm, wo) * co < R(0))) return V3<R>(0, 0, 0);
        R F = FresnelDielectric(Fabs(Dot(wo, wm)), ior);
        R T = R(1) - F;
        if (refl) { R M = MicrofacetBRDF(wo, wi, wm, ggx); R v = F * M; return V3<R>(v, v, v); }
        R temp = Dot(wi, wm) * eta + Dot(wo, wm);
        R dw = Fabs(Dot(wi, wm)) * Fabs(Dot(wo, wm)) / (temp * temp);
        R Tss = ggx.D(wm) * ggx.G(wo, wi) * dw / (Fabs(ci * co));
        return T * albedo * Tss;
    }
    R pdfTransparent(V3<R> wo, V3<R> wi) const {
        R co = wo.z, ci = wi.z;
        bool refl = co * ci > R(0);
        R eta = R(1);
        if (!refl) eta = (!backface) ? ior : R(1) / ior;
        V3<R> wm = (eta * wi + wo);
        wm = Normalize((wm.z > R(0)) ? wm : -wm);
        if (Dot(wm, wi) * ci < R(0) || Dot(wm, wo) * co < R(0)) return R(0);
        R F = FresnelDielectric(Fabs(Dot(wo, wm)), ior);
        R T = R(1) - F;
        if (refl) return pdfGlossy(wo, wi, wm) * F;
        R temp = Dot(wi, wm) + Dot(wo, wm) / eta;
        R dw = Fabs(Dot(wo, wm)) / (temp * temp);
        R noh = Fmax(wm.z, R(0));
        R D = ggx.D(wm);
        return D * noh * dw * T;
    }
    BSDFSample<R> sampleOpaque(V3<R> wo, V2<R> u, R X) const {
        const R gp = R(0.5f);
        if (X < gp) {
            V3<R> wm = ggx.sampleWm(wo, u);
            V3<R> wi = Reflect(-wo, wm);
            V3<R> f = fGlossy(wo, wi, wm);
            f = f * (R(1) / gp);
            R pdf = pdfGlossy(wo, wi, wm);
            return {Glossy | Reflected, f, wi, pdf};
        }
        // sampleCosineWeightedHemisphere, samplers.cuh:88-95
        R phi = R(2) * R(PI_F) * u.x;
        R ct = Sqrt(u.y);
        R st = Sqrt(R(1) - u.y);
        V3<R> wi(st * Mt<R>::cos(phi), st * Mt<R>::sin(phi), ct);
        V3<R> f = albedo * fDiffuse(wo, wi);
        f = f * (R(1) / (R(1) - gp));
        R pdf = pdfDiffuse(wo, wi);
        return {Diffuse | Reflected, f, wi, pdf};
    }
    BSDFSample<R> sampleConductor(V3<R> wo, V2<R> u) const {
        V3<R> wm = ggx.sampleWm(wo, u);
        V3<R> wi = Reflect(-wo, wm);
        V3<R> f = fConductor(wo, wi);
        R pdf = pdfGlossy(wo, wi, wm);
        return {Glossy | Reflected, f, wi, pdf};
    }
    BSDFSample<R> sampleTransparent(V3<R> wo, V2<R> u, R X) const {
        const bool inside = backface;
        R eta = inside ? (R(1) / ior) : ior;
        V3<R> ht = ggx.sampleWm(wo, u);
        V3<R> wi;
        bool tir = !Refract(wo, ht, eta, &wi);
        R gp = R(0.5f);
        if ((X < gp && !inside) || (tir && inside)) {
            wi = Reflect(-wo, ht);
            if (!SameHemi(wi, wo, V3<R>(0, 0, 1))) return {Absorbed, V3<R>(0, 0, 0), V3<R>(0, 0, 0), R(0)};
            R M = MicrofacetBRDF(wo, wi, ht, ggx);
            R F = FresnelDielectric(Dot(wo, ht), eta);
            R v = F * M;
            V3<R> f(v, v, v);
            if (!tir) f = f * (R(1) / gp);
            R pdf = pdfGlossy(wo, wi, ht);
            return {Reflected | Glossy, f, wi, pdf};
        }
        if (tir || SameHemi(wi, wo, V3<R>(0, 0, 1)) || wi.z == R(0)) return {Absorbed, V3<R>(0, 0, 0), V3<R>(0, 0, 0), R(0)};
        R pdf = pdfBTDF(wo, wi, ht, eta);
        V3<R> f = fBTDF(wo, wi, ht, eta);
        if (!inside) f = f * (R(1) / (R(1) - gp));
        return {Transmitted | Glossy, f, wi, pdf};
    }
    V3<R> f(V3<R> wwo, V3<R> wwi) const {
        M3<R> ib = Transpose3(basis);
        V3<R> wo = Mul3(ib, wwo);
        V3<R> wi = Mul3(ib, wwi);
        R wm = metallic;
        R wt = (R(1) - wm) * transmission;
        R wd = (R(1) - wm) * (R(1) - transmission);
        V3<R> F(0, 0, 0);
        if (wm > R(0)) F = F + (fConductor(wo, wi) * wm);
        if (wt > R(0)) F = F + (fTransparent(wo, wi) * wt);
        if (wd > R(0)) F = F + (fOpaque(wo, wi) * wd);
        return F;
    }
    R pdf(V3<R> wwo, V3<R> wwi) const {
        M3<R> ib = Transpose3(basis);
        V3<R> wi = Mul3(ib, wwi);
        V3<R> wo = Mul3(ib, wwo);
        R wm = metallic;
        R wt = (R(1) - wm) * transmission;
        R wd = (R(1) - wm) * (R(1) - transmission);
        R p = 0;
        if (wm > R(0)) p += wm * pdfConductor(wo, wi);
        if (wt > R(0)) p += wt * pdfTransparent(wo, wi);
        if (wd > R(0)) p += wd * pdfOpaque(wo, wi);
        return p;
    }
    BSDFSample<R> sampleF(V3<R> wwo, V2<R> u, V2<R> X2) const {
        R pp = X2.x;
        V3<R> wo = Mul3(Transpose3(basis), wwo);
        R pm = metallic;
        R pt = (R(1) - pm) * transmission;
        BSDFSample<R> bs;
        if (pp < pm) bs = sampleConductor(wo, u);
        else if (pp < pt) bs = sampleTransparent(wo, u, X2.y);  // non-cumulative (bsdf.cuh:256)
        else bs = sampleOpaque(wo, u, X2.y);
        bs.wi = Mul3(basis, bs.wi);
        return bs;
    }
    void regularize() {
        if (roughness < R(0.3f)) roughness = Clamp(R(2) * roughness, R(0.1f), R(0.3f));
        ggx.regularize();
    }
    bool nonSpecular() const { return !ggx.smooth(); }
};

// ------------------------------------------------------------------ sky
// intersectSphere with the centre at the origin, atmosphere.cuh:30-57
template <class R> bool IntersectSphere(const Ray<R>& r, R radius, R* t0, R* t1) {
    V3<R> oc = V3<R>(0, 0, 0) - r.o;
    R a = Dot(r.d, r.d);
    R b = R(-2) * Dot(r.d, oc);
    R c = Dot(oc, oc) - Sqr(radius);
    R disc = Sqr(b) - R(4) * a * c;
    if (disc < R(0)) return false;
    R sq = Sqrt(disc);
    *t0 = (-b - sq) / (R(2) * a);
    *t1 = (-b + sq) / (R(2) * a);
    if (*t0 > *t1) std::swap(*t0, *t1);
    if (*t1 < R(0)) return false;
    return true;
}
// NishitaAtmosphereModel::sampleLi, atmosphere.cuh:132-223
template <class R> V3<R> SkyLi(const Ray<R>& ray, V3<R> sun, R emission) {
    const R eR = R(EARTH_R_F), aR = R(ATMOS_R_F), Hr = R(HR_F), Hm = R(HM_F);
    const V3<R> sR(R(3.8e-6f), R(13.5e-6f), R(33.1e-6f));
    const V3<R> sM(R(21e-6f), R(21e-6f), R(21e-6f));
    const R aRay = R(0.0f), aMie = R(21e-6f * 0.1f);
    R tmin = 0, tmax = Inf<R>();
    {
        R te, tx;
        if (!IntersectSphere(ray, aR, &te, &tx) || tx < R(0)) return V3<R>(0, 0, 0);
        if (te > tmin && te > R(0)) tmin = te;
        if (tx < tmax) tmax = tx;
    }
    const R depth = tmax - tmin;
    const R step = depth / R(16);
    R cur = tmin;
    V3<R> iR(0, 0, 0), iM(0, 0, 0);
    R pRsum = 0, pMsum = 0;
    for (uint32_t i = 0; i < 16; ++i) {
        const V3<R> sp = ray.at(cur + (step * R(0.5f)));
        const R alt = Length(sp) - eR;
        const R pR = Mt<R>::exp(-alt / Hr) * step;
        const R pM = Mt<R>::exp(-alt / Hm) * step;
        pRsum += pR;
        pMsum += pM;
        R le = 0, lx = 0;
        Ray<R> rl(sp, sun);
        IntersectSphere(rl, aR, &le, &lx);
        const R lstep = lx / R(8);
        R curl = 0, lR = 0, lM = 0;
        uint32_t j;
        for (j = 0; j < 8; ++j) {
            const V3<R> spl = rl.at(curl + (lstep * R(0.5f)));
            const R altl = Length(spl) - eR;
            if (altl < R(0)) break;
            lR += Mt<R>::exp(-altl / Hr) * lstep;
            lM += Mt<R>::exp(-altl / Hm) * lstep;
            curl += lstep;
        }
        if (j == 8) {
            const V3<R> tR = sR + aRay;
            const V3<R> tM = sM + aMie;
            const V3<R> tau = tR * (pRsum + lR) + tM * (pMsum + lM);
            const V3<R> Tr(Mt<R>::exp(-tau.x), Mt<R>::exp(-tau.y), Mt<R>::exp(-tau.z));
            iR = iR + Tr * pR;
            iM = iM + Tr * pM;
        }
        cur += step;
    }
    const R mu = Dot(ray.d, sun);
    const R ray_phase = (R(3) * (R(1) + Sqr(mu))) / (R(16.0f * PI_F));
    const R g = R(0.76f);
    const R denom = R(8.0f * PI_F) * ((R(2) + Sqr(g)) * Mt<R>::pow(R(1) + Sqr(g) - R(2) * g * mu, R(1.5f)));
    const R mie_phase = R(3) * ((R(1) - Sqr(g)) * (R(1) + Sqr(mu))) / denom;
    const V3<R> L = (sR * iR * ray_phase + sM * iM * mie_phase);
    return L * R(SUN_NITS_F) * emission;
}

// ------------------------------------------------------------------ integrator
struct FrameConst {
    const orc_scene* S;
    kpt_settings st;
    M4<float> inv_view, inv_proj;
    V3<float> cam_pos;
    M4<float> curr_vp, prev_vp;
    // DebugData (blas.cuh:39-42): when set, the node pops of every
    // closest-hit trace of a path are added here (shadow rays use a dummy,
    // integrator.cuh:129); entry_bound selects the GPU kernels' BLAS entry.
    Counters* path_counts = nullptr;
    bool entry_bound = false;
};

template <class R> struct SurfInt {  // interaction.cuh:15-67
    int32_t prim = -1, inst = -1;
    V3<R> wo, pos, n;
    const Light* light = nullptr;
    int32_t material = -1;
    R distance = Inf<R>();
    V2<R> uv;
    bool backface = false;
};

template <class R> struct VisSurf {  // packing.cuh:156-176
    V3<R> albedo;
    R depth = Inf<R>();
    V3<R> normal, position;
    int32_t inst = -1, prim = -1;
};

// Texture::evaluate, texture.cuh:34-66
template <class R> V3<R> TexEval(const orc_scene* S, int tex, V2<R> uv) {
    const Texture& t = S->textures[tex];
    if (t.pixel_buffer_index < 0 || t.pixel_buffer_index >= (int64_t)S->pixels.size()) return V3<R>(0, 0, 0);
    size_t pxs = (size_t)t.channel_count * (t.bit_depth / 8);
    size_t rows = (size_t)t.width * pxs;
    V2<R> st;
    st.x = uv.x - Floor(uv.x);
    st.y = uv.y - Floor(uv.y);
    int32_t x = int32_t(st.x * R(t.width)) % t.width;
    int32_t y = int32_t(st.y * R(t.height)) % t.height;
    if (x < 0) x += t.width;
    if (y < 0) y += t.height;
    size_t off = (size_t)y * rows + (size_t)x * pxs;
    if (off + 3 >= S->pixels.size()) return V3<R>(0, 0, 0);
    const uint8_t* p = &S->pixels[t.pixel_buffer_index + off];
    return V3<R>(R((float)(int)p[0] * INV_255_F), R((float)(int)p[1] * INV_255_F), R((float)(int)p[2] * INV_255_F));
}
// sRGB decode, color.cuh:428-433
template <class R> R SrgbDecode(R v) {
    const R U = R(0.04045f), A = R(12.92f), C = R(0.055f), T = R(2.4f);
    return (v <= U) ? (v / A) : Mt<R>::pow((v + C) / (R(1) + C), T);
}

// Material::getBSDF, material.cu:17-70
template <class R> BSDF<R> GetBSDF(const orc_scene* S, const SurfInt<R>& si) {
    const Material& m = S->materials[si.material];
    V3<R> albedo = cvt<R>(m.albedo);
    R rough = R(m.roughness), metal = R(m.metallic), trans = R(m.transmission);
    V3<R> wg = si.n;
    if (m.albedo_texture_id >= 0) {
        V3<R> s = TexEval(S, m.albedo_texture_id, si.uv);
        V3<R> lin(SrgbDecode(s.x), SrgbDecode(s.y), SrgbDecode(s.z));
        albedo = albedo * lin;
    }
    if (m.orm_texture_id >= 0) {
        V3<R> s = TexEval(S, m.orm_texture_id, si.uv);
        rough *= s.y;
        metal *= s.z;
    }
    if (m.transmission_texture_id >= 0) {
        V3<R> s = TexEval(S, m.transmission_texture_id, si.uv);
        trans *= s.x;
    }
    if (m.normal_texture_id >= 0) {
        V3<R> s = TexEval(S, m.normal_texture_id, si.uv);
        V3<R> mn = (s * R(2)) - R(1);
        mn.x *= R(m.normal_scale);
        mn.y *= R(m.normal_scale);
        M3<R> fr = OnbFrisvad(wg);
        V3<R> sn = Mul3(fr, Normalize(mn));
        wg = (Dot(si.wo, sn) > R(0)) ? sn : wg;
    }
    BSDF<R> b;
    b.basis = OnbFrisvad(wg);
    b.albedo = albedo;
    b.metallic = metal;
    b.roughness = rough;
    b.transmission = trans;
    b.ior = R(m.ior);
    b.backface = si.backface;
    b.ggx.alpha = Sqr(rough);
    return b;
}

// Light::L, light.cu:20-31
template <class R> V3<R> LightL(const orc_scene* S, const Light& l, V2<R> uv) {
    const Material& m = S->materials[S->tris[l.prim_id].material_id];
    V3<R> e = cvt<R>(m.emissive_factor * m.emission_scale_nits);
    if (m.emission_texture_id >= 0) {
        V3<R> s = TexEval(S, m.emission_texture_id, uv);
        V3<R> lin(Mt<R>::pow(s.x, R(2.2f)), Mt<R>::pow(s.y, R(2.2f)), Mt<R>::pow(s.z, R(2.2f)));
        e = e * lin;
    }
    return e;
}
// Light::pdf_Li, light.cu:11-18
template <class R> R LightPdfLi(const Light& l, V3<R> ctx_pos, V3<R> lpos, V3<R> lnrm) {
    R area_pdf = R(1) / R(l.area);
    R dist = Distance(lpos, ctx_pos);
    R c = AbsDot(Normalize(lpos - ctx_pos), lnrm);
    return area_pdf / (c / Sqr(dist));
}

// Intersection::getSurfaceInteraction, interaction.cu:11-38
template <class R> SurfInt<R> GetSurf(const orc_scene* S, const Isect<R>& it, const Ray<R>& ray) {
    SurfInt<R> si;
    si.prim = it.prim; si.inst = it.inst; si.distance = it.distance;
    si.wo = -ray.d;
    const Triangle& tri = S->tris[it.prim];
    M4<R> model = CvtM4<R>(Inverse(S->meshes[it.inst].curr_inv_model));
    si.material = tri.material_id;
    si.pos = ray.at(it.distance);
    si.n = Normalize(xyz(Mul(model, V4<R>(cvt<R>(tri.local_geometric_normal), R(0)))));
    V2<R> uv;
    uv.x = (it.bary.x * R(tri.t[0].x)) + (it.bary.y * R(tri.t[1].x)) + (it.bary.z * R(tri.t[2].x));
    uv.y = (it.bary.x * R(tri.t[0].y)) + (it.bary.y * R(tri.t[1].y)) + (it.bary.z * R(tri.t[2].y));
    si.uv = uv;
    if (!SameHemi(si.n, si.wo)) { si.n = si.n * R(-1); si.backface = true; }
    if (tri.light_id >= 0) si.light = &S->lights[tri.light_id];
    return si;
}

// Light::sampleLi with Triangle::sample (light.cu:33-45, triangle.cuh:104-119):
// a point of the light's triangle from two uniform numbers (the square-to-
// triangle map of triangle.cuh:104-119), the light's world normal, the pdf
// of that point in solid angle seen from `from`, and the emitted radiance.
template <class R> struct LightSample { V3<R> p, n; R pdf = 0; V3<R> L; V2<R> uv; };
template <class R> LightSample<R> SampleLight(const orc_scene* S, const Light& l, V3<R> from, V2<R> u2) {
    const Triangle& tri = S->tris[l.prim_id];
    M4<R> model = CvtM4<R>(Inverse(S->meshes[l.inst_id].curr_inv_model));
    V2<R> bb;
    if (u2.x < u2.y) { bb.x = u2.x / R(2); bb.y = u2.y - bb.x; }
    else { bb.y = u2.y / R(2); bb.x = u2.x - bb.y; }
    V3<R> bary(bb.x, bb.y, R(1) - bb.x - bb.y);
    LightSample<R> ls;
    ls.uv.x = (bary.x * R(tri.t[0].x)) + (bary.y * R(tri.t[1].x)) + (bary.z * R(tri.t[2].x));
    ls.uv.y = (bary.x * R(tri.t[0].y)) + (bary.y * R(tri.t[1].y)) + (bary.z * R(tri.t[2].y));
    V3<R> p = (cvt<R>(tri.p[0]) * bary.x) + (cvt<R>(tri.p[1]) * bary.y) + (cvt<R>(tri.p[2]) * bary.z);
    ls.p = xyz(Mul(model, V4<R>(p, R(1))));
    ls.n = xyz(Mul(model, V4<R>(cvt<R>(tri.local_geometric_normal), R(0))));
    R area = R(0.5f) * Length(Cross(cvt<R>(tri.p[1]) - cvt<R>(tri.p[0]), cvt<R>(tri.p[2]) - cvt<R>(tri.p[0])));
    R spdf = R(1) / area;
    spdf *= Sqr(Distance(ls.p, from));
    spdf /= AbsDot(Normalize(ls.p - from), ls.n);
    ls.pdf = spdf;
    ls.L = LightL(S, l, ls.uv);
    return ls;
}

template <class R> bool Unoccluded(const orc_scene* S, const SurfInt<R>& si, V3<R> target) {  // integrator.cuh:124-131
    V3<R> org = si.backface ? si.pos - (si.n * R(HIT_EPS_F)) : si.pos + (si.n * R(HIT_EPS_F));
    Ray<R> sr(org, Normalize(target - org));
    R tmax = Distance(target, sr.o) - R(0.11f);
    Counters cn;
    const bool occluded = TlasIntersectP(S, sr, R(0), tmax, &cn);
    AddEvents(cn);
    return !occluded;
}

template <class R> V3<R> SunDir(const kpt_settings& st) {  // integrator.cuh:222-230
    R phi = R(st.sun_phi_rad), th = R(st.sun_theta_rad);
    return Normalize(V3<R>(Mt<R>::cos(phi) * Mt<R>::cos(th), Mt<R>::sin(th), Mt<R>::sin(phi) * Mt<R>::cos(th)));
}

// Integrator::Li, integrator.cuh:241-402
// A path that ends in the radiance cache at its bounce-1 vertex: the query
// the renderer evaluates there (position, regularised roughness, outgoing
// direction, throughput times the vertex's reflectance proxy 0.04 + 0.96
// albedo); `hit` is false for paths that never reach that vertex (or whose
// depth limit is 1).  `nee` (input): the vertex's next-event estimate is made
// before the path ends; otherwise the cache stands for it too.
struct CacheQuery {
    bool nee = false;
    float wide_roughness = 0.0f;   // input: lobes this rough (or diffuse samples) lead to the cache vertex
    bool hit = false;
    float pos[3], rough, dir[3], beta[3];
};

template <class R> V3<R> Li(const FrameConst& fc, Ray<R> ray, Sampler<R>& sampler, VisSurf<R>* vs,
                            CacheQuery* cq = nullptr) {
    const orc_scene* S = fc.S;
    const int32_t max_depth = fc.st.max_depth;
    V3<R> L(0, 0, 0), beta(1, 1, 1);
    bool any_nonspec = false;
    R eta_scale = R(1.0);
    int32_t depth = 0;
    R p_b = R(1);
    V3<R> prev_pos(0, 0, 0);
    bool wide_in = false;   // the current vertex was reached through a wide lobe (cache queries only)
    const V3<R> sun = SunDir<R>(fc.st);
    const R emis = R(fc.st.sun_emission_factor);
    const V3<R> observer(0, R(EARTH_R_F + 1), 0);
    const int nlights = (int)S->lights.size();
    while (!IsZero(beta)) {
        SetPhase<R>(depth == 0 ? PH_SHADE0 : PH_SHADE);
        sampler.seed = sampler.seed + (uint32_t)depth;
        bool first = (depth == 0);
        Counters cn;
        Isect<R> it = TlasIntersect(S, ray, R(0), Inf<R>(), &cn, fc.entry_bound);
        AddEvents(cn);
        if (fc.path_counts) {   // dbg.blas_hits / dbg.tlas_hits (blas.cuh:140, tlas.cu:87)
            fc.path_counts->interior += cn.interior;
            fc.path_counts->leaf += cn.leaf;
            fc.path_counts->tlas += cn.tlas;
        }
        if (it.prim < 0) {
            SetPhase<R>(first ? PH_SKY0 : PH_SKY);
            V3<R> sky = SkyLi(Ray<R>(observer, ray.d), sun, emis);
            if (first) {
                // LeSun, integrator.cuh:135-143
                R min_sim = Mt<R>::cos(R(fc.st.sun_angular_diameter_rad) / R(2));
                R sim = Dot(ray.d, sun);
                bool on_sun = sim > min_sim;
                R sr = emis * R(SUN_NITS_F);
                R le = sr * Mt<R>::pow(R(1) / sr, on_sun ? R(0) : R(1));
                sky = sky * V3<R>(le, le, le);
                vs->albedo = sky; vs->depth = Inf<R>(); vs->normal = V3<R>(0, 0, 0); vs->position = V3<R>(0, 0, 0);
                vs->inst = -1; vs->prim = -1;
            }
            L = L + sky * beta;
            break;
        }
        SurfInt<R> si = GetSurf(S, it, ray);
        V3<R> Le(0, 0, 0);
        if (si.light) Le = LightL(S, *si.light, si.uv);
        if (!IsZero(Le)) {
            R wl = R(1);
            if (si.light && !first) {
                R lpdf = (R(1) / R((float)nlights)) * LightPdfLi(*si.light, prev_pos, si.pos, si.n);
                R f = R(1) * p_b, g = R(1) * lpdf;
                wl = Sqr(f) / (Sqr(f) + Sqr(g));
            }
            L = L + Le * wl * beta;
        }
        BSDF<R> bsdf = GetBSDF(S, si);
        if (first) {
            V3<R> alb = !IsZero(Le) ? Le : bsdf.albedo;
            vs->albedo = alb; vs->depth = si.distance; vs->normal = si.n; vs->position = si.pos;
            vs->inst = si.inst; vs->prim = si.prim;
            bsdf.albedo = V3<R>(1, 1, 1);
        }
        if (any_nonspec) bsdf.regularize();
        const bool to_cache = cq && depth >= 1 && wide_in && depth < max_depth;
        auto record_query = [&]() {
            cq->hit = true;
            cq->pos[0] = (float)si.pos.x; cq->pos[1] = (float)si.pos.y; cq->pos[2] = (float)si.pos.z;
            cq->rough = (float)bsdf.roughness;
            cq->dir[0] = (float)-ray.d.x; cq->dir[1] = (float)-ray.d.y; cq->dir[2] = (float)-ray.d.z;
            // throughput times the vertex's reflectance proxy (0.04 + 0.96 albedo)
            V3<R> bq = beta * (V3<R>(R(0.04f), R(0.04f), R(0.04f)) + bsdf.albedo * R(0.96f));
            cq->beta[0] = (float)bq.x; cq->beta[1] = (float)bq.y; cq->beta[2] = (float)bq.z;
        };
        if (to_cache && !cq->nee) {   // the cache stands for all radiance scattered here
            record_query();
            break;
        }
        if (bsdf.nonSpecular()) {
            // the renderer's cache vertex samples only the sun when there are
            // no area lights; every other vertex keeps the reference's 0.5
            const R sun_p = (to_cache && nlights == 0) ? R(1.0f) : R(0.5f);
            if (sampler.get1D() < sun_p) {
                // sampleLdSun, integrator.cuh:147-186
                V3<R> Ld(0, 0, 0);
                V3<R> sun_pos = sun * R(SOLAR_DIST_F);
                R sun_radius = (R(2) * R(SOLAR_DIST_F) * Mt<R>::tan(R(fc.st.sun_angular_diameter_rad) / R(2))) / R(2);
                V2<R> u2 = sampler.get2D();
                R u1 = sampler.get1D();
                V3<R> off(u2.x * R(2) - R(1), u2.y * R(2) - R(1), u1 * R(2) - R(1));
                V3<R> target = sun_pos + (off * sun_radius);
                V3<R> wo = -ray.d;
                V3<R> fcos = bsdf.f(wo, sun) * Fmax(R(0), Dot(sun, (si.backface ? R(-1) : R(1)) * si.n));
                bool ok = !IsZero(fcos);
                V3<R> srad(0, 0, 0);
                if (ok) {
                    // the sky toward the sun: the same for every call of a frame
                    PhaseScope<R> phase_(PH_SKYSUN);
                    srad = SkyLi(Ray<R>(observer, sun), sun, emis);
                    ok = !IsZero(srad);
                }
                if (ok) ok = Unoccluded(S, si, target);
                if (ok) {
                    R sun_area = R(PI_F) * Sqr(sun_radius);
                    V3<R> snrm = Normalize(target - sun_pos), wi = Normalize(target - si.pos);
                    R cth = AbsDot(snrm, -wi);
                    R pdf = (R(1) / sun_area) / (cth / Sqr(R(SOLAR_DIST_F)));
                    Ld = (fcos * srad * R(SUN_NITS_F)) / pdf;
                }
                L = L + Ld * beta / sun_p;
            } else {
                // sampleLd, integrator.cuh:188-220
                V3<R> Ld(0, 0, 0);
                R ul = sampler.get1D();
                if (nlights >= 1) {
                    int li = std::min(int(ul * R((float)nlights)), nlights - 1);
                    const Light& l = S->lights[li];
                    R pmf = R(1) / R((float)nlights);
                    V2<R> u2 = sampler.get2D();
                    LightSample<R> ls = SampleLight(S, l, si.pos, u2);
                    V3<R> p = ls.p;
                    R spdf = ls.pdf;
                    V3<R> wi = Normalize(p - si.pos);
                    V3<R> lL = ls.L;
                    bool ok = !(spdf <= R(0) || IsZero(lL));
                    V3<R> wo = -ray.d;
                    V3<R> fcos(0, 0, 0);
                    if (ok) { fcos = bsdf.f(wo, wi) * Fmax(R(0), Dot(wi, (si.backface ? R(-1) : R(1)) * si.n)); ok = !IsZero(fcos); }
                    if (ok) ok = Unoccluded(S, si, p);
                    if (ok) {
                        R pl = pmf * spdf;
                        R pb = bsdf.pdf(wo, wi);
                        R f = R(1) * pl, g = R(1) * pb;
                        R wl = Sqr(f) / (Sqr(f) + Sqr(g));
                        Ld = (lL * fcos * wl) / pl;
                    }
                }
                L = L + Ld * beta / (R(1) - sun_p);
            }
        }
        if (to_cache) {   // the cache completes what the continuation of this path would add
            record_query();
            break;
        }
        if (depth++ == max_depth) break;
        V3<R> wo = -ray.d;
        V2<R> u2 = sampler.get2D();
        V2<R> X2 = sampler.get2D();
        BSDFSample<R> bs = bsdf.sampleF(wo, u2, X2);
        if (IsZero(bs.f)) break;
        beta = beta * (bs.f * AbsDot(si.n, bs.wi) / bs.pdf);
        p_b = bsdf.pdf(wo, bs.wi);
        bool spec = (bs.scatter & Specular) != 0;
        any_nonspec |= !spec;
        wide_in = ((bs.scatter & Diffuse) != 0) || (cq && bsdf.roughness >= R(cq->wide_roughness));
        prev_pos = si.pos;
        // SurfaceInteraction::spawnRay, interaction.cuh:31-41
        V3<R> org = (bs.scatter & Transmitted) ? si.pos - (si.n * R(HIT_EPS_F)) : si.pos + (si.n * R(HIT_EPS_F));
        ray = Ray<R>(org, bs.wi);
        // russianRoulette, sampling.cuh:19-31
        V3<R> rr = beta * eta_scale;
        if (MaxComp(rr) < R(1) && depth > 1) {
            R q = Fmax(R(0), R(1) - MaxComp(rr));
            if (sampler.get1D() < q) break;
            beta = beta / (R(1) - q);
        }
    }
    return L;
}

// Camera::generateRay, camera.cuh:16-26, from getShadingJob (shading_kernel.cuh:15-27)
template <class R> Ray<R> CameraRay(const FrameConst& fc, int px, int py, int W, int H) {
    float u = ((float)px + 0.5f) / (float)W;
    float v = ((float)py + 0.5f) / (float)H;
    R nx = R(2.0f * u - 1.0f), ny = R(2.0f * v - 1.0f);
    M4<R> ip = CvtM4<R>(fc.inv_proj), iv = CvtM4<R>(fc.inv_view);
    V4<R> tcs = Mul(ip, V4<R>(nx, ny, R(1), R(1)));
    V4<R> tws = Mul(iv, V4<R>(Normalize(xyz(tcs) / tcs.w), R(0)));
    V3<R> dir = Normalize(xyz(tws));
    return Ray<R>(cvt<R>(fc.cam_pos), dir);
}

// GBuffer packing, packing.cuh:13-17, 105-125, 178-197
inline uint32_t FloatToByte(float v) { v = Saturate(v); return uint32_t(v * 255.0f) & 0xFFu; }
inline void EncodeUnit(V3<float> v, float* ex, float* ey) {
    float ax = std::fabs(v.x), ay = std::fabs(v.y), az = std::fabs(v.z);
    float sum = ax + ay + az;
    float x = v.x / sum, y = v.y / sum;
    if (v.z <= 0.0f) {
        float rx = 1.0f - std::fabs(y);
        float ry = 1.0f - std::fabs(x);
        x = x >= 0.0f ? rx : -rx;
        y = y >= 0.0f ? ry : -ry;
    }
    *ex = x; *ey = y;
}

// computeVelocity, kernels.cu:165-194 (camera products are per-frame constants)
inline V2<float> Velocity(const FrameConst& fc, const VisSurf<float>& vs) {
    V2<float> vel;
    if (vs.inst < 0) return vel;
    const TriangleMesh& m = fc.S->meshes[vs.inst];
    M4<float> prev_model = Inverse(m.prev_inv_model);
    V3<float> local = xyz(Mul(m.curr_inv_model, V4<float>(vs.position, 1.0f)));
    V4<float> cc = Mul(fc.curr_vp, V4<float>(vs.position, 1.0f));
    V4<float> pc = Mul(MulMM(fc.prev_vp, prev_model), V4<float>(local, 1.0f));
    V3<float> cn = xyz(cc) / cc.w, pn = xyz(pc) / pc.w;
    float cux = (cn.x + 1.0f) / 2.0f, cuy = (cn.y + 1.0f) / 2.0f;
    float pux = (pn.x + 1.0f) / 2.0f, puy = (pn.y + 1.0f) / 2.0f;
    vel.x = cux - pux; vel.y = cuy - puy;
    return vel;
}

FrameConst MakeFrame(const orc_scene* S, const kpt_settings* st, const kpt_camera* cam) {
    FrameConst fc;
    fc.S = S;
    fc.st = *st;
    fc.inv_view = LoadM4(cam->inv_view);
    fc.inv_proj = LoadM4(cam->inv_proj);
    fc.cam_pos = xyz(fc.inv_view.c[3]);
    fc.curr_vp = MulMM(Inverse(fc.inv_proj), Inverse(fc.inv_view));
    fc.prev_vp = MulMM(Inverse(LoadM4(cam->prev_inv_proj)), Inverse(LoadM4(cam->prev_inv_view)));
    return fc;
}

// clampOutput, color.cuh:295-302
template <class R> V3<R> ClampOutput(V3<R> v) {
    if (IsNan(v.x) || IsNan(v.y) || IsNan(v.z) || IsInf(v.x) || IsInf(v.y) || IsInf(v.z)) return V3<R>(0, 0, 0);
    const R mx = R(1.0e8f);
    return V3<R>(Clamp(v.x, R(0), mx), Clamp(v.y, R(0), mx), Clamp(v.z, R(0), mx));
}

template <class Fn> void ParallelFor(int64_t n, int threads, Fn fn) {
    if (threads <= 1 || n < 64) { for (int64_t i = 0; i < n; ++i) fn(i); return; }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            for (;;) {
                int64_t b = next.fetch_add(16);
                if (b >= n) break;
                int64_t e = std::min<int64_t>(n, b + 16);
                for (int64_t i = b; i < e; ++i) fn(i);
            }
        });
    for (auto& th : pool) th.join();
}

template <class R>
void RenderPixel(const FrameConst& fc, int W, int H, int pidx, int frame, V3<R>* sample, VisSurf<R>* vs) {
    int px = pidx % W, py = pidx / W;
    Sampler<R> sm;
    sm.initPixelSeed(px, py, W, frame, fc.st.seed_mode);
    // BoxFilter::sample draws two values whose jitter is then discarded (kernels.cu:397-402)
    (void)sm.get2D();
    SetPhase<R>(PH_OTHER);
    Ray<R> ray = CameraRay<R>(fc, px, py, W, H);
    *vs = VisSurf<R>();
    V3<R> L = Li<R>(fc, ray, sm, vs);
    SetPhase<R>(PH_OTHER);
    // RGBSpectrum(fs.weight) * camera_weight * Li: 1 * 1 * L (kernels.cu:411-412)
    V3<R> w(R(1.0f), R(1.0f), R(1.0f));
    L = (w * R(1.0f)) * L;
    *sample = ClampOutput(L);
}

// testReprojection / testReprojectedDepth (kernels.cu:196-222)
inline bool TestReprojection(const FrameConst& fc, const VisSurf<float>& vs, float ppx, float ppy,
                             const uint32_t* prev_g, int W, int H, V3<float> prev_cam) {
    const TriangleMesh& m = fc.S->meshes[vs.inst];
    V3<float> local = xyz(Mul(m.curr_inv_model, V4<float>(vs.position, 1.0f)));
    M4<float> prev_model = Inverse(m.prev_inv_model);
    int ix = (int)ppx, iy = (int)ppy;
    ix = ix < 0 ? 0 : (ix > W - 1 ? W - 1 : ix);
    iy = iy < 0 ? 0 : (iy > H - 1 ? H - 1 : iy);
    float prev_depth = kpt::dm::u2f(prev_g[4 * ((size_t)iy * W + ix) + 3]);
    V3<float> prev_world = xyz(Mul(prev_model, V4<float>(local, 1.0f)));
    float d = Length(prev_cam - prev_world);
    return std::fabs(d - prev_depth) < (prev_depth * 0.045f);
}

// reprojectAccumulate (kernels.cu:299-338) without SUPER_SAMPLE
inline void ReprojectAccumulate(const FrameConst& fc, int W, int H, int px, int py, const VisSurf<float>& vs,
                                V2<float> vel, V3<float> curr, const float* prev_acc, const uint32_t* prev_g,
                                V3<float> prev_cam, float out[4]) {
    auto pass = [&](float w) { out[0] = curr.x; out[1] = curr.y; out[2] = curr.z; out[3] = w; };
    if (vs.inst < 0) return pass(0.0f);
    float ppx = (float)px - vel.x * (float)W, ppy = (float)py - vel.y * (float)H;
    if (ppx < 0 || ppx >= (float)W || ppy < 0 || ppy >= (float)H) return pass(0.0f);
    if (!TestReprojection(fc, vs, ppx, ppy, prev_g, W, H, prev_cam)) return pass(0.0f);
    // DeviceTextureBuffer::textureReadBilinear(p, false): alpha of the first tap
    int dx = (int)ppx, dy = (int)ppy;
    float ws = ppx - (float)dx, wt = ppy - (float)dy;
    auto tap = [&](int x, int y, int c) {
        x = x < 0 ? 0 : (x > W - 1 ? W - 1 : x);
        y = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
        return prev_acc[4 * ((size_t)y * W + x) + c];
    };
    float prev[3];
    for (int c = 0; c < 3; ++c) {
        float t0 = tap(dx, dy, c) + ws * (tap(dx + 1, dy, c) - tap(dx, dy, c));
        float t1 = tap(dx, dy + 1, c) + ws * (tap(dx + 1, dy + 1, c) - tap(dx, dy + 1, c));
        prev[c] = t0 + wt * (t1 - t0);
    }
    float history = tap(dx, dy, 3);
    float alpha = 1.0f / std::fmin(history + 1, 16.0f);
    const float cur[3] = {curr.x, curr.y, curr.z};
    for (int c = 0; c < 3; ++c) out[c] = prev[c] + alpha * (cur[c] - prev[c]);
    out[3] = history + 1;
}

}  // namespace orc

using namespace orc;

extern "C" {

orc_scene* orc_scene_create(const kpt_scene_desc* d) {
    orc_scene* S = new orc_scene();
    const int32_t T = d->num_triangles;
    // Instances: the reference has one per mesh (renderer.cu:386-417).
    int32_t NI = d->num_instances > 0 ? d->num_instances : d->num_meshes;
    std::vector<int32_t> inst_mesh(NI);
    std::vector<M4<float>> inst_model(NI);
    for (int32_t i = 0; i < NI; ++i) {
        inst_mesh[i] = d->num_instances > 0 ? d->instance_mesh[i] : i;
        inst_model[i] = LoadM4(d->num_instances > 0 ? d->instance_model + 16 * i : d->mesh_model + 16 * i);
    }
    for (int32_t i = 0; i < d->num_materials; ++i) {
        const kpt_material& m = d->materials[i];
        Material o;
        o.albedo_texture_id = m.albedo_texture_id;
        o.albedo = V3<float>(m.albedo[0], m.albedo[1], m.albedo[2]);
        o.orm_texture_id = m.orm_texture_id;
        o.metallic = m.metallic; o.roughness = m.roughness;
        o.transmission_texture_id = m.transmission_texture_id; o.transmission = m.transmission;
        o.ior = m.ior;
        o.emission_texture_id = m.emission_texture_id;
        o.emissive_factor = V3<float>(m.emission_factor[0], m.emission_factor[1], m.emission_factor[2]);
        o.emission_scale_nits = m.emission_scale_nits;
        o.normal_texture_id = m.normal_texture_id; o.normal_scale = m.normal_scale;
        S->materials.push_back(o);
    }
    for (int32_t i = 0; i < d->num_textures; ++i) {
        Texture t;
        t.width = (int)d->textures[i].width; t.height = (int)d->textures[i].height;
        t.channel_count = (int)d->textures[i].channels; t.bit_depth = 8;
        t.pixel_buffer_index = (int64_t)d->textures[i].offset;
        S->textures.push_back(t);
    }
    S->pixels.assign(d->pixels, d->pixels + d->num_pixel_bytes);
    // Triangle constructor, triangle.cuh:54-64
    S->tris.resize(T);
    for (int32_t t = 0; t < T; ++t) {
        Triangle& tr = S->tris[t];
        for (int v = 0; v < 3; ++v) {
            const float* p = d->positions + 9 * (size_t)t + 3 * v;
            const float* n = d->normals + 9 * (size_t)t + 3 * v;
            const float* uv = d->uvs + 6 * (size_t)t + 2 * v;
            tr.p[v] = V3<float>(p[0], p[1], p[2]);
            tr.n[v] = V3<float>(n[0], n[1], n[2]);
            tr.t[v].x = uv[0]; tr.t[v].y = uv[1];
        }
        tr.material_id = d->material_ids[t];
        V3<float> e0 = tr.p[1] - tr.p[0], e1 = tr.p[2] - tr.p[0];
        V3<float> gn = Cross(e0, e1);
        V3<float> avg = (tr.n[0] + tr.n[1] + tr.n[2]) / 3.0f;
        float dd = Dot(gn, avg);
        V3<float> lg = (dd < 0.0f) ? -gn : gn;
        tr.local_geometric_normal = Normalize(lg);
    }
    // Emissive triangles become lights, renderer.cu:391-401 (one per instance use).
    std::vector<int32_t> first_inst_of_mesh(d->num_meshes, -1);
    for (int32_t i = 0; i < NI; ++i) if (first_inst_of_mesh[inst_mesh[i]] < 0) first_inst_of_mesh[inst_mesh[i]] = i;
    for (int32_t m = 0; m < d->num_meshes; ++m) {
        for (int32_t t = d->mesh_first[m]; t < d->mesh_first[m] + d->mesh_count[m]; ++t) {
            const Material& mat = S->materials[S->tris[t].material_id];
            V3<float> e = mat.emissive_factor * mat.emission_scale_nits;
            S->tris[t].mesh_id = first_inst_of_mesh[m];
            if (e.x > 0.0f || e.y > 0.0f || e.z > 0.0f) {
                Light l;
                V3<float> c = Cross(S->tris[t].p[1] - S->tris[t].p[0], S->tris[t].p[2] - S->tris[t].p[0]);
                l.area = 0.5f * Length(c);
                l.prim_id = t;
                l.inst_id = first_inst_of_mesh[m];
                S->tris[t].light_id = (int32_t)S->lights.size();
                S->lights.push_back(l);
            }
        }
    }
    // BLAS per geometry mesh (BLASBuilder::build, as_builder.cuh:26-35)
    BlasBuilder bb;
    bb.S = S;
    bb.cache.resize(T);
    std::vector<Bounds3f> geom_bounds(d->num_meshes);
    for (int32_t m = 0; m < d->num_meshes; ++m) {
        S->geom_first.push_back(d->mesh_first[m]);
        S->geom_count.push_back(d->mesh_count[m]);
        int md = 0;
        int64_t root = d->mesh_count[m] > 0 ? bb.build(d->mesh_first[m], d->mesh_count[m], &geom_bounds[m], &md) : -1;
        S->geom_root.push_back(root);
        S->max_blas_depth = std::max(S->max_blas_depth, md);
    }
    // Instances: BLAS2 + TriangleMesh, then setMeshTransform (renderer.cu:250-269)
    for (int32_t i = 0; i < NI; ++i) {
        int32_t m = inst_mesh[i];
        BLAS2 b;
        b.root = S->geom_root[m];
        b.bounds = geom_bounds[m];
        b.original_bounds = geom_bounds[m];
        b.mesh_id = i;
        BlasSetTransform(b, inst_model[i]);
        S->blas.push_back(b);
        TriangleMesh tm;
        tm.prim_offset = d->mesh_first[m];
        tm.prim_count = d->mesh_count[m];
        tm.blas_index = i;
        tm.curr_inv_model = GlmInverse(inst_model[i]);
        tm.prev_inv_model = tm.curr_inv_model;
        S->meshes.push_back(tm);
        S->inst_geom.push_back(m);
    }
    BuildTlas(S);
    return S;
}

void orc_scene_destroy(orc_scene* s) { delete s; }

void orc_scene_info(const orc_scene* s, int64_t out[8]) {
    out[0] = (int64_t)s->tris.size();
    out[1] = (int64_t)s->nodes.size();
    out[2] = (int64_t)s->tlas_nodes.size();
    out[3] = (int64_t)s->blas.size();
    out[4] = s->max_blas_depth;
    out[5] = s->max_tlas_depth;
    out[6] = (int64_t)s->lights.size();
    out[7] = (int64_t)s->materials.size();
}

int64_t orc_get_blas(const orc_scene* s, int32_t mesh, void* nodes32, int64_t cap, int32_t* tri_index, int64_t cap_tris) {
    if (mesh < 0 || mesh >= (int32_t)s->geom_root.size() || s->geom_root[mesh] < 0) return -1;
    int64_t root = s->geom_root[mesh];
    int64_t count = 2 * (int64_t)s->geom_count[mesh] - 1;
    if (count > cap || s->geom_count[mesh] > cap_tris) return -2;
    BVH2Node* out = (BVH2Node*)nodes32;
    for (int64_t i = 0; i < count; ++i) {
        BVH2Node n = s->nodes[root + i];
        if (n.isLeaf()) n.first -= s->geom_first[mesh];                 // leaf: slot relative to the mesh
        else if (n.first >= 0) n.first -= (int32_t)root;                // interior: child relative to the root
        out[i] = n;
    }
    // tri_index slice for this mesh: BLASes are appended in mesh order
    int64_t ti0 = 0;
    for (int32_t m = 0; m < mesh; ++m) ti0 += s->geom_count[m];
    for (int32_t i = 0; i < s->geom_count[mesh]; ++i) tri_index[i] = s->tri_index[ti0 + i];
    return count;
}

int64_t orc_get_tlas(const orc_scene* s, void* nodes32, int64_t cap) {
    int64_t n = (int64_t)s->tlas_nodes.size();
    if (n > cap) return -2;
    std::memcpy(nodes32, s->tlas_nodes.data(), n * sizeof(TLASNode));
    return n;
}

void orc_trace(const orc_scene* s, const kpt_ray* rays, int64_t n, int32_t mode, kpt_hit* hits,
               int32_t* occluded, kpt_trace_counters* counters, int32_t* max_stack, int32_t threads) {
    std::vector<Counters> per(std::max(1, threads));
    std::atomic<int> ms{0};
    std::vector<Counters> pc(n);
    ParallelFor(n, threads, [&](int64_t i) {
        const kpt_ray& r = rays[i];
        Ray<float> ray(V3<float>(r.org[0], r.org[1], r.org[2]), V3<float>(r.dir[0], r.dir[1], r.dir[2]));
        Counters cn;
        if (mode == 0 || mode == 2 || mode == 4) {
            Isect<float> it = mode == 2 ? BruteClosest(s, ray, r.tmin, r.tmax)
                                        : TlasIntersect(s, ray, r.tmin, r.tmax, &cn, mode == 4);
            hits[i].t = it.distance; hits[i].prim_id = it.prim; hits[i].inst_id = it.inst;
            hits[i].bary[0] = it.bary.x; hits[i].bary[1] = it.bary.y; hits[i].bary[2] = it.bary.z;
        } else {
            bool o = mode == 1 ? TlasIntersectP(s, ray, r.tmin, r.tmax, &cn) : BruteAny(s, ray, r.tmin, r.tmax);
            occluded[i] = o ? 1 : 0;
        }
        pc[i] = cn;
    });
    if (counters) {
        kpt_trace_counters c{};
        int mx = 0;
        for (int64_t i = 0; i < n; ++i) {
            c.blas_interior_pops += pc[i].interior; c.blas_leaf_pops += pc[i].leaf; c.tri_tests += pc[i].tri;
            c.blas_entries += pc[i].entries; c.tlas_pops += pc[i].tlas;
            mx = std::max(mx, pc[i].max_stack);
        }
        c.rays = (uint64_t)n;
        *counters = c;
        if (max_stack) *max_stack = mx;
    }
    (void)ms; (void)per;
}

void orc_render_pixels(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam, int32_t W, int32_t H,
                       const int32_t* pixels, int64_t n, int32_t frame0, int32_t spp, int32_t dp, float* out_rgb,
                       uint32_t* out_g, float* out_v, int32_t threads) {
    FrameConst fc = MakeFrame(s, settings, cam);
    ParallelFor(n, threads, [&](int64_t i) {
        int p = pixels[i];
        if (dp) {
            V3<double> acc(0, 0, 0);
            for (int f = 0; f < spp; ++f) {
                V3<double> smp; VisSurf<double> vs;
                RenderPixel<double>(fc, W, H, p, frame0 + f, &smp, &vs);
                acc = acc + smp;
            }
            acc = acc / double(spp);
            out_rgb[3 * i] = (float)acc.x; out_rgb[3 * i + 1] = (float)acc.y; out_rgb[3 * i + 2] = (float)acc.z;
            return;
        }
        // Integrator::addSample (integrator.cuh:406-416): running sum in the
        // accumulation buffer, estimate = sum / float(frame_index + 1).
        V3<float> acc(0, 0, 0), est(0, 0, 0);
        VisSurf<float> vs;
        for (int f = 0; f < spp; ++f) {
            V3<float> smp;
            RenderPixel<float>(fc, W, H, p, frame0 + f, &smp, &vs);
            acc = acc + smp;
            est = acc / float(frame0 + f + 1);
        }
        out_rgb[3 * i] = est.x; out_rgb[3 * i + 1] = est.y; out_rgb[3 * i + 2] = est.z;
        if (out_g) {
            float ex, ey;
            EncodeUnit(vs.normal, &ex, &ey);
            uint32_t alb = FloatToByte(vs.albedo.x) | (FloatToByte(vs.albedo.y) << 8) | (FloatToByte(vs.albedo.z) << 16);
            out_g[4 * i] = kpt::dm::f2u(ex); out_g[4 * i + 1] = kpt::dm::f2u(ey);
            out_g[4 * i + 2] = alb; out_g[4 * i + 3] = kpt::dm::f2u(vs.depth);
        }
        if (out_v) {
            V2<float> vel = Velocity(fc, vs);
            out_v[4 * i] = vel.x; out_v[4 * i + 1] = vel.y;
            out_v[4 * i + 2] = (float)vs.prim; out_v[4 * i + 3] = (float)vs.inst;
        }
    });
}

// One frame in which every path ends in the radiance cache at its bounce-1
// vertex (no training paths): per pixel the radiance gathered before the
// cache takes over (up to and including that vertex's next-event estimate
// when `nee`; unclamped; out_L n x 3) and the cache query
// (out_q n x 12: hit, position xyz, roughness, direction xyz, throughput xyz,
// unused).
void orc_render_cache_queries(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam, int32_t W,
                              int32_t H, int32_t frame, int32_t nee, float wide_roughness, float* out_L, float* out_q,
                              int32_t threads) {
    FrameConst fc = MakeFrame(s, settings, cam);
    ParallelFor((int64_t)W * H, threads, [&](int64_t i) {
        int px = (int)(i % W), py = (int)(i / W);
        Sampler<float> sm;
        sm.initPixelSeed(px, py, W, frame, fc.st.seed_mode);
        (void)sm.get2D();
        Ray<float> ray = CameraRay<float>(fc, px, py, W, H);
        VisSurf<float> vs;
        CacheQuery cq;
        cq.nee = nee != 0;
        cq.wide_roughness = wide_roughness;
        V3<float> L = Li<float>(fc, ray, sm, &vs, &cq);
        V3<float> w(1.0f, 1.0f, 1.0f);
        L = (w * 1.0f) * L;
        out_L[3 * i] = L.x; out_L[3 * i + 1] = L.y; out_L[3 * i + 2] = L.z;
        float* q = out_q + 12 * i;
        q[0] = cq.hit ? 1.0f : 0.0f;
        for (int k = 0; k < 3; ++k) { q[1 + k] = cq.pos[k]; q[5 + k] = cq.dir[k]; q[8 + k] = cq.beta[k]; }
        q[4] = cq.rough;
        q[11] = 0.0f;
        if (!cq.hit) for (int k = 1; k < 12; ++k) q[k] = 0.0f;
    });
}

// Renderer::setMeshTransform (renderer.cu:250-269): TriangleMesh::setTransform
// keeps the previous inverse (triangle.cuh:142-146), BLAS2::setTransform
// (blas.cuh:190-200); then the per-frame TLAS rebuild (renderer.cu:115).
void orc_set_instance_transform(orc_scene* s, int32_t i, const float model[16]) {
    M4<float> M = LoadM4(model);
    BlasSetTransform(s->blas[i], M);
    TriangleMesh& tm = s->meshes[i];
    tm.prev_inv_model = tm.curr_inv_model;
    tm.curr_inv_model = GlmInverse(M);
    BuildTlas(s);
}

// One frame with integrator_use_temporal_accumulation (kernels.cu:385-451):
// every pixel's sample, G/V-buffers and the reprojected render value from the
// previous frame's accumulation and G-buffer.
}  // extern "C"

namespace orc {
// One pixel of a temporal-accumulation frame: the sample, its G/V-buffer
// words and reprojectAccumulate against the previous frame's buffers.
inline void TemporalPixel(const FrameConst& fc, int W, int H, int pidx, int frame, const float* prev_accum,
                          const uint32_t* prev_gbuffer, V3<float> prev_cam, float* out_render, uint32_t* out_g,
                          float* out_v) {
    V3<float> smp;
    VisSurf<float> vs;
    RenderPixel<float>(fc, W, H, pidx, frame, &smp, &vs);
    float ex, ey;
    EncodeUnit(vs.normal, &ex, &ey);
    uint32_t alb = FloatToByte(vs.albedo.x) | (FloatToByte(vs.albedo.y) << 8) | (FloatToByte(vs.albedo.z) << 16);
    out_g[0] = kpt::dm::f2u(ex); out_g[1] = kpt::dm::f2u(ey);
    out_g[2] = alb; out_g[3] = kpt::dm::f2u(vs.depth);
    V2<float> vel = Velocity(fc, vs);
    out_v[0] = vel.x; out_v[1] = vel.y;
    out_v[2] = (float)vs.prim; out_v[3] = (float)vs.inst;
    ReprojectAccumulate(fc, W, H, pidx % W, pidx / W, vs, vel, smp, prev_accum, prev_gbuffer, prev_cam, out_render);
}
}  // namespace orc

extern "C" {

void orc_render_temporal(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam, int32_t W,
                         int32_t H, int32_t frame, const float* prev_accum, const uint32_t* prev_gbuffer,
                         float* out_render, uint32_t* out_g, float* out_v, int32_t threads) {
    FrameConst fc = MakeFrame(s, settings, cam);
    V3<float> prev_cam = xyz(LoadM4(cam->prev_inv_view).c[3]);
    ParallelFor((int64_t)W * H, threads, [&](int64_t i) {
        TemporalPixel(fc, W, H, (int)i, frame, prev_accum, prev_gbuffer, prev_cam, out_render + 4 * i, out_g + 4 * i,
                      out_v + 4 * i);
    });
}

void orc_render_temporal_pixels(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam, int32_t W,
                                int32_t H, int32_t frame, const float* prev_accum, const uint32_t* prev_gbuffer,
                                const int32_t* pixels, int64_t n, float* out_render, uint32_t* out_g, float* out_v,
                                int32_t threads) {
    FrameConst fc = MakeFrame(s, settings, cam);
    V3<float> prev_cam = xyz(LoadM4(cam->prev_inv_view).c[3]);
    ParallelFor(n, threads, [&](int64_t i) {
        TemporalPixel(fc, W, H, pixels[i], frame, prev_accum, prev_gbuffer, prev_cam, out_render + 4 * i,
                      out_g + 4 * i, out_v + 4 * i);
    });
}

void orc_sample_pixels(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam, int32_t W, int32_t H,
                       const int32_t* pixels, int64_t n, int32_t frame, float* out_rgb, int32_t threads) {
    FrameConst fc = MakeFrame(s, settings, cam);
    ParallelFor(n, threads, [&](int64_t i) {
        V3<float> smp; VisSurf<float> vs;
        RenderPixel<float>(fc, W, H, pixels[i], frame, &smp, &vs);
        out_rgb[3 * i] = smp.x; out_rgb[3 * i + 1] = smp.y; out_rgb[3 * i + 2] = smp.z;
    });
}

void orc_dmath(int32_t fn, const float* x, const float* y, float* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        float a = x[i];
        switch (fn) {
            case 0: out[i] = kpt::dm::expf(a); break;
            case 1: out[i] = kpt::dm::exp2f(a); break;
            case 2: out[i] = kpt::dm::log2f(a); break;
            case 3: out[i] = kpt::dm::powf(a, y[i]); break;
            case 4: out[i] = kpt::dm::sinf(a); break;
            case 5: out[i] = kpt::dm::cosf(a); break;
            case 6: out[i] = kpt::dm::tanf(a); break;
            case 7: out[i] = kpt::dm::acosf(a); break;
            case 8: out[i] = kpt::dm::logf(a); break;
            default: out[i] = 0.0f;
        }
    }
}

uint32_t orc_pcg_hash(uint32_t v) { return PcgHash(v); }

}  // extern "C"

namespace orc {
template <class R> BSDF<R> ProbeBSDF(const float* p) {
    BSDF<R> b;
    b.basis = M3<R>();
    b.albedo = V3<R>(R(p[0]), R(p[1]), R(p[2]));
    b.metallic = R(p[3]);
    b.roughness = R(p[4]);
    b.transmission = R(p[5]);
    b.ior = R(p[6]);
    b.backface = p[7] != 0.0f;
    b.ggx.alpha = Sqr(b.roughness);
    return b;
}
template <class R> void BsdfEval(const float* params, const float* wo, const float* wi, int64_t n, float* of, float* op) {
    BSDF<R> b = ProbeBSDF<R>(params);
    for (int64_t i = 0; i < n; ++i) {
        V3<R> o{R(wo[3 * i]), R(wo[3 * i + 1]), R(wo[3 * i + 2])}, w{R(wi[3 * i]), R(wi[3 * i + 1]), R(wi[3 * i + 2])};
        V3<R> f = b.f(o, w);
        of[3 * i] = (float)f.x; of[3 * i + 1] = (float)f.y; of[3 * i + 2] = (float)f.z;
        op[i] = (float)b.pdf(o, w);
    }
}
template <class R> void BsdfSample(const float* params, const float* wo, const float* u, int64_t n, float* owi, float* of,
                                   float* op, int32_t* ofl) {
    BSDF<R> b = ProbeBSDF<R>(params);
    for (int64_t i = 0; i < n; ++i) {
        V3<R> o{R(wo[3 * i]), R(wo[3 * i + 1]), R(wo[3 * i + 2])};
        V2<R> u2, x2;
        u2.x = R(u[4 * i]); u2.y = R(u[4 * i + 1]); x2.x = R(u[4 * i + 2]); x2.y = R(u[4 * i + 3]);
        BSDFSample<R> s = b.sampleF(o, u2, x2);
        owi[3 * i] = (float)s.wi.x; owi[3 * i + 1] = (float)s.wi.y; owi[3 * i + 2] = (float)s.wi.z;
        of[3 * i] = (float)s.f.x; of[3 * i + 1] = (float)s.f.y; of[3 * i + 2] = (float)s.f.z;
        op[i] = (float)s.pdf;
        ofl[i] = s.scatter;
    }
}
}  // namespace orc

extern "C" {

void orc_bsdf_eval(const float* params, const float* wo, const float* wi, int64_t n, int32_t dp, float* out_f,
                   float* out_pdf) {
    if (dp) BsdfEval<double>(params, wo, wi, n, out_f, out_pdf);
    else BsdfEval<float>(params, wo, wi, n, out_f, out_pdf);
}

void orc_bsdf_sample(const float* params, const float* wo, const float* u, int64_t n, int32_t dp, float* out_wi,
                     float* out_f, float* out_pdf, int32_t* out_flags) {
    if (dp) BsdfSample<double>(params, wo, u, n, out_wi, out_f, out_pdf, out_flags);
    else BsdfSample<float>(params, wo, u, n, out_wi, out_f, out_pdf, out_flags);
}

void orc_ggx_d(float alpha, const float* wm, int64_t n, int32_t dp, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        if (dp) { GGX<double> g; g.alpha = alpha; out[i] = (float)g.D(V3<double>(wm[3 * i], wm[3 * i + 1], wm[3 * i + 2])); }
        else { GGX<float> g; g.alpha = alpha; out[i] = g.D(V3<float>(wm[3 * i], wm[3 * i + 1], wm[3 * i + 2])); }
    }
}

void orc_fresnel_dielectric(const float* c, const float* ior, int64_t n, int32_t dp, float* out) {
    for (int64_t i = 0; i < n; ++i)
        out[i] = dp ? (float)FresnelDielectric<double>(c[i], ior[i]) : FresnelDielectric<float>(c[i], ior[i]);
}

void orc_sky_li(const float* o, const float* d, int64_t n, const float sun[3], float emission, int32_t dp, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        if (dp) {
            Ray<double> r(V3<double>(o[3 * i], o[3 * i + 1], o[3 * i + 2]), V3<double>(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
            V3<double> L = SkyLi(r, V3<double>(sun[0], sun[1], sun[2]), (double)emission);
            out[3 * i] = (float)L.x; out[3 * i + 1] = (float)L.y; out[3 * i + 2] = (float)L.z;
        } else {
            Ray<float> r(V3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]), V3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
            V3<float> L = SkyLi(r, V3<float>(sun[0], sun[1], sun[2]), emission);
            out[3 * i] = L.x; out[3 * i + 1] = L.y; out[3 * i + 2] = L.z;
        }
    }
}

void orc_camera_rays(const kpt_camera* cam, int32_t W, int32_t H, const int32_t* pixels, int64_t n, float* org,
                     float* dir) {
    kpt_settings st;
    std::memset(&st, 0, sizeof(st));
    FrameConst fc = MakeFrame(nullptr, &st, cam);
    for (int64_t i = 0; i < n; ++i) {
        Ray<float> r = CameraRay<float>(fc, pixels[i] % W, pixels[i] / W, W, H);
        org[3 * i] = r.o.x; org[3 * i + 1] = r.o.y; org[3 * i + 2] = r.o.z;
        dir[3 * i] = r.d.x; dir[3 * i + 1] = r.d.y; dir[3 * i + 2] = r.d.z;
    }
}

void orc_glm_inverse(const float m[16], float out[16]) { StoreM4(GlmInverse(LoadM4(m)), out); }
void orc_mat4_inverse(const float m[16], float out[16]) { StoreM4(Inverse(LoadM4(m)), out); }

}  // extern "C"

// ------------------------------------------------------------------ shading probes
// Per-routine views of the integrator's building blocks for the pinning
// tests; each calls the same functions Li calls.
extern "C" {

/* getSurfaceInteraction (interaction.cu:11-38) and Material::getBSDF
 * (material.cu:17-70, texture.cuh:34-66) at each ray's closest hit.
 * out [n][24]: hit, t, position xyz, world geometric normal xyz (flipped to
 * face -d), uv, backface, material, albedo rgb, metallic, roughness,
 * transmission, ior, shading normal xyz (the BSDF basis' third column),
 * light index (-1 none), 0. */
void orc_shade_probe(const orc_scene* s, const kpt_ray* rays, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        float* o = out + 24 * i;
        std::memset(o, 0, 24 * sizeof(float));
        Ray<float> r(V3<float>(rays[i].org[0], rays[i].org[1], rays[i].org[2]),
                     V3<float>(rays[i].dir[0], rays[i].dir[1], rays[i].dir[2]));
        Counters cn;
        Isect<float> it = TlasIntersect(s, r, rays[i].tmin, rays[i].tmax, &cn);
        o[22] = -1.0f;
        if (it.prim < 0) continue;
        SurfInt<float> si = GetSurf(s, it, r);
        BSDF<float> b = GetBSDF(s, si);
        o[0] = 1.0f; o[1] = si.distance;
        o[2] = si.pos.x; o[3] = si.pos.y; o[4] = si.pos.z;
        o[5] = si.n.x; o[6] = si.n.y; o[7] = si.n.z;
        o[8] = si.uv.x; o[9] = si.uv.y;
        o[10] = si.backface ? 1.0f : 0.0f;
        o[11] = (float)si.material;
        o[12] = b.albedo.x; o[13] = b.albedo.y; o[14] = b.albedo.z;
        o[15] = b.metallic; o[16] = b.roughness; o[17] = b.transmission; o[18] = b.ior;
        V3<float> z = Mul3(b.basis, V3<float>(0.0f, 0.0f, 1.0f));
        o[19] = z.x; o[20] = z.y; o[21] = z.z;
        o[22] = si.light ? (float)(si.light - s->lights.data()) : -1.0f;
    }
}

/* Light::sampleLi (light.cu:33-45) of light `light` seen from positions
 * from[n][3] with numbers u[n][2]: out [n][12] = point xyz, world normal
 * xyz (unnormalised, as the reference transforms it), pdf in solid angle,
 * emitted radiance rgb, uv.  pdf_li [n] (optional): Light::pdf_Li
 * (light.cu:11-18) of the same point, the pdf the emitter-hit MIS weight
 * uses. */
void orc_light_sample(const orc_scene* s, int32_t light, const float* from, const float* u, int64_t n, float* out,
                      float* pdf_li) {
    const Light& l = s->lights[light];
    for (int64_t i = 0; i < n; ++i) {
        V3<float> f(from[3 * i], from[3 * i + 1], from[3 * i + 2]);
        V2<float> u2;
        u2.x = u[2 * i]; u2.y = u[2 * i + 1];
        LightSample<float> ls = SampleLight(s, l, f, u2);
        float* o = out + 12 * i;
        o[0] = ls.p.x; o[1] = ls.p.y; o[2] = ls.p.z;
        o[3] = ls.n.x; o[4] = ls.n.y; o[5] = ls.n.z;
        o[6] = ls.pdf;
        o[7] = ls.L.x; o[8] = ls.L.y; o[9] = ls.L.z;
        o[10] = ls.uv.x; o[11] = ls.uv.y;
        if (pdf_li) pdf_li[i] = LightPdfLi(l, f, ls.p, Normalize(ls.n));
    }
}

/* encodeUnitVector (packing.cuh:105-125): out [n][2]. */
void orc_encode_unit(const float* v, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) EncodeUnit(V3<float>(v[3 * i], v[3 * i + 1], v[3 * i + 2]), &out[2 * i], &out[2 * i + 1]);
}

/* floatToByte (packing.cuh:13-17). */
void orc_float_to_byte(const float* v, int64_t n, uint32_t* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = FloatToByte(v[i]);
}

}  // extern "C"

extern "C" {

/* Algorithmic FP32 op counts of one sample of each listed pixel (frame
 * `frame`): the oracle instantiated with the counting float CF.
 * out[phase * 4 + kind], phases: 0 traversal, 1 sky of primary misses,
 * 2 sky of later misses, 3 sky toward the sun (a per-frame constant the
 * renderer evaluates once), 4 shading at the primary vertex, 5 shading at
 * later vertices, 6 camera rays and output; kinds: 0 + - * /, 1 sqrt,
 * 2 transcendentals, 3 min / max / abs / floor.  Single-threaded. */
void orc_count_ops(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam, int32_t W, int32_t H,
                   const int32_t* pixels, int64_t n, int32_t frame, uint64_t* out) {
    FrameConst fc = MakeFrame(s, settings, cam);
    OpCounts c;
    std::memset(&c, 0, sizeof(c));
    g_ops = &c;
    for (int64_t i = 0; i < n; ++i) {
        V3<CF> smp;
        VisSurf<CF> vs;
        RenderPixel<CF>(fc, W, H, pixels[i], frame, &smp, &vs);
    }
    g_ops = nullptr;
    g_phase = PH_OTHER;
    std::memcpy(out, &c, sizeof(c));
}

// Per-pixel DebugData of one sample (kernels.cu:404-416, the commented-out
// heat map at :442-443): out[2i] = BLAS node pops (blas_hits), out[2i+1] =
// TLAS node pops (tlas_hits) of pixel pixels[i]'s closest-hit traces in frame
// `frame`.  entry_bound: traversal as the GPU kernels enter BLASes.
void orc_heat_map(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam, int32_t W, int32_t H,
                  const int32_t* pixels, int64_t n, int32_t frame, int32_t entry_bound, int64_t* out) {
    FrameConst fc = MakeFrame(s, settings, cam);
    fc.entry_bound = entry_bound != 0;
    for (int64_t i = 0; i < n; ++i) {
        Counters cn;
        fc.path_counts = &cn;
        V3<float> smp;
        VisSurf<float> vs;
        RenderPixel<float>(fc, W, H, pixels[i], frame, &smp, &vs);
        out[2 * i] = (int64_t)(cn.interior + cn.leaf);
        out[2 * i + 1] = (int64_t)cn.tlas;
    }
}

// Traversal events of one sample of each pixel pixels[i] in frame `frame`
// (every closest-hit and shadow ray of its path, traversed as the reference
// does): out[6i .. 6i+5] = BLAS interior pops, BLAS leaf pops, triangle tests,
// BLAS entries, TLAS pops, rays traced.
void orc_sample_events(const orc_scene* s, const kpt_settings* settings, const kpt_camera* cam, int32_t W, int32_t H,
                       const int32_t* pixels, int64_t n, int32_t frame, int32_t threads, int64_t* out) {
    FrameConst fc = MakeFrame(s, settings, cam);
    ParallelFor(n, threads, [&](int64_t i) {
        SampleEvents ev;
        g_events = &ev;
        V3<float> smp;
        VisSurf<float> vs;
        RenderPixel<float>(fc, W, H, pixels[i], frame, &smp, &vs);
        g_events = nullptr;
        const int64_t v[6] = {(int64_t)ev.interior, (int64_t)ev.leaf, (int64_t)ev.tri, (int64_t)ev.entries,
                              (int64_t)ev.tlas, (int64_t)ev.rays};
        std::memcpy(out + 6 * i, v, sizeof(v));
    });
}

}  // extern "C"

