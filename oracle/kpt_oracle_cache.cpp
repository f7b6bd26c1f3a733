// kpt_oracle_cache.cpp — CPU oracle of the radiance-cache network: input
// encoding, forward, relative-L2 loss, backward and the Adam step.
// TEST INFRASTRUCTURE ONLY: tests/ call it as the reference the GPU kernels are compared with
// (kittlespt_amd/csrc/hip/kpt_cache.hip); the product never links it.
//
// The reference renderer has no radiance cache; the definition followed here
// is the network's own specification: 32 features (sin/cos of the position
// normalised to a box at 1, 2, 4, 8 cycles; sin/cos of (pi/2) * direction;
// roughness; 1) -> 64 -> 64 -> 64 -> 64 -> 64 -> 3, ReLU on hidden layers, no
// biases; loss (o - t)^2 / (o^2 + 0.01) per channel with the denominator held
// constant, averaged over samples and channels; Adam (0.9, 0.99, 1e-15).
//
// Two arithmetic modes:
//   mode 0  mirrors the GPU's rounding points: weights, features, hidden
//           activations and back-propagated gradients are rounded to BF16;
//           sums are accumulated in double (the GPU accumulates in FP32 in an
//           unspecified order inside the matrix cores, so the comparison has
//           a tolerance).
//   mode 1  exact FP64 throughout (no rounding); pins the backward pass by
//           finite differences.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "kpt_oracle.h"
#include "../kittlespt_amd/csrc/common/kpt_dmath.h"

namespace orcc {

constexpr int NL = 6;
constexpr int IN[NL] = {32, 64, 64, 64, 64, 64};
constexpr int OUT[NL] = {64, 64, 64, 64, 64, 3};
constexpr int OFF[NL] = {0, 2048, 6144, 10240, 14336, 18432};
constexpr int NP = 18624;

inline float bf(float x) {  // round to BF16 (nearest even), back to float
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) {
        u = (u | 0x400000u) & 0xffff0000u;
    } else {
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

struct Enc {
    float bmin[3], binv[3], scale;
};

// encoding of one query: pos = (x, y, z, roughness), dir = (x, y, z, -)
void Encode(const float* pos, const float* dir, const Enc& e, float* f) {
    const float PI = 3.14159265358979323846f, HALF_PI = 1.57079632679489661923f;
    for (int k = 0; k < 4; ++k)
        for (int a = 0; a < 3; ++a) {
            const float ph = (pos[a] - e.bmin[a]) * e.binv[a];
            const float arg = (PI * (float)(1 << k)) * ph;
            f[6 * k + 2 * a] = kpt::dm::sinf(arg);
            f[6 * k + 2 * a + 1] = kpt::dm::cosf(arg);
        }
    for (int a = 0; a < 3; ++a) {
        const float arg = HALF_PI * dir[a];
        f[24 + 2 * a] = kpt::dm::sinf(arg);
        f[25 + 2 * a] = kpt::dm::cosf(arg);
    }
    f[30] = pos[3];
    f[31] = 1.0f;
}

// forward of one sample; acts[L] = input of layer L (after rounding), out = 3 outputs
void Forward(const double* w, const float* feat, int mode, std::vector<double>* acts, double out[3]) {
    std::vector<double> x(feat, feat + 32);
    if (mode == 0)
        for (double& v : x) v = bf((float)v);
    for (int L = 0; L < NL; ++L) {
        acts[L] = x;
        std::vector<double> y(OUT[L]);
        for (int o = 0; o < OUT[L]; ++o) {
            double s = 0.0;
            for (int i = 0; i < IN[L]; ++i) s += w[OFF[L] + o * IN[L] + i] * x[i];
            y[o] = s;
        }
        if (L == NL - 1) {
            for (int c = 0; c < 3; ++c) out[c] = mode == 0 ? (double)(float)y[c] : y[c];
            return;
        }
        for (double& v : y) {
            v = v > 0.0 ? v : 0.0;
            if (mode == 0) v = bf((float)v);
        }
        x.swap(y);
    }
}

std::vector<double> Weights(const float* w, int mode) {
    std::vector<double> d(NP);
    for (int i = 0; i < NP; ++i) d[i] = mode == 0 ? bf(w[i]) : w[i];
    return d;
}

}  // namespace orcc

using namespace orcc;

extern "C" {

void orc_cache_encode(const float* pos, const float* dir, int64_t n, const float enc[7], float* out) {
    Enc e;
    std::memcpy(e.bmin, enc, 12);
    std::memcpy(e.binv, enc + 3, 12);
    e.scale = enc[6];
    for (int64_t i = 0; i < n; ++i) Encode(pos + 4 * i, dir + 4 * i, e, out + 32 * i);
}

void orc_cache_forward(const float* w, const float* pos, const float* dir, int64_t n, const float enc[7],
                       int32_t mode, double* out) {
    Enc e;
    std::memcpy(e.bmin, enc, 12);
    std::memcpy(e.binv, enc + 3, 12);
    e.scale = enc[6];
    std::vector<double> wd = Weights(w, mode);
    std::vector<double> acts[NL];
    float f[32];
    for (int64_t i = 0; i < n; ++i) {
        Encode(pos + 4 * i, dir + 4 * i, e, f);
        Forward(wd.data(), f, mode, acts, out + 3 * i);
    }
}

// Loss and its gradient over n samples.  w_den (optional) supplies the
// weights whose outputs form the loss denominators (finite-difference checks
// of the stop-gradient loss); grad (optional) receives dLoss/dw.
double orc_cache_loss_grad(const float* w, const float* w_den, const float* pos, const float* dir, const float* tgt,
                           int64_t n, const float enc[7], int32_t mode, double* grad) {
    Enc e;
    std::memcpy(e.bmin, enc, 12);
    std::memcpy(e.binv, enc + 3, 12);
    e.scale = enc[6];
    std::vector<double> wd = Weights(w, mode);
    std::vector<double> wdd = w_den ? Weights(w_den, mode) : wd;
    if (grad) std::memset(grad, 0, sizeof(double) * NP);
    const double inv_norm_d = 1.0 / (3.0 * (double)n);
    const float inv_norm_f = 1.0f / (3.0f * (float)n);
    double loss = 0.0;
    std::vector<double> acts[NL], acts_den[NL];
    float f[32];
    for (int64_t s = 0; s < n; ++s) {
        Encode(pos + 4 * s, dir + 4 * s, e, f);
        double o[3], od[3];
        Forward(wd.data(), f, mode, acts, o);
        if (w_den) Forward(wdd.data(), f, mode, acts_den, od);
        else for (int c = 0; c < 3; ++c) od[c] = o[c];
        std::vector<double> dy(3);
        for (int c = 0; c < 3; ++c) {
            const double t = tgt[4 * s + c];
            if (mode == 0) {
                const float of = (float)o[c], d = of - (float)t, den = of * of + 0.01f;
                loss += (double)(d * d / den);
                dy[c] = bf((2.0f * d / den) * inv_norm_f);
            } else {
                const double d = o[c] - t, den = od[c] * od[c] + 0.01;
                loss += d * d / den;
                dy[c] = 2.0 * d / den * inv_norm_d;
            }
        }
        if (!grad) continue;
        for (int L = NL - 1; L >= 0; --L) {
            const std::vector<double>& x = acts[L];
            for (int oo = 0; oo < OUT[L]; ++oo)
                for (int i = 0; i < IN[L]; ++i) grad[OFF[L] + oo * IN[L] + i] += dy[oo] * x[i];
            if (L == 0) break;
            std::vector<double> dx(IN[L], 0.0);
            for (int i = 0; i < IN[L]; ++i) {
                double sum = 0.0;
                for (int oo = 0; oo < OUT[L]; ++oo) sum += wd[OFF[L] + oo * IN[L] + i] * dy[oo];
                double v = x[i] > 0.0 ? sum : 0.0;
                if (mode == 0) v = bf((float)v);
                dx[i] = v;
            }
            dy.swap(dx);
        }
    }
    return mode == 0 ? loss / (3.0 * (double)n) : loss * inv_norm_d;
}

// The Adam update the GPU applies, in float with the same expressions.
void orc_cache_adam(float* w, float* m, float* v, const float* grad, int64_t step, float lr) {
    const float b1 = 0.9f, b2 = 0.99f, eps = 1e-15f;
    const float bc1 = (float)(1.0 - std::pow(0.9, (double)step));
    const float bc2 = (float)(1.0 - std::pow(0.99, (double)step));
    for (int i = 0; i < NP; ++i) {
        const float g = grad[i];
        m[i] = b1 * m[i] + (1.0f - b1) * g;
        v[i] = b2 * v[i] + (1.0f - b2) * (g * g);
        const float mh = m[i] / bc1, vh = v[i] / bc2;
        w[i] = w[i] - lr * mh / (std::sqrt(vh) + eps);
    }
}

}  // extern "C"
