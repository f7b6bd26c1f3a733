// kpt_oracle_post.cpp — CPU oracle of the frame's post chain: albedo
// modulation, log-luminance histogram and auto exposure, bloom mip chain,
// exposure in Yxy and the AgX film transform.  TEST INFRASTRUCTURE ONLY: tests/
// call it as the reference to compare with; the product never links it.
//
// Scalar float code in IEEE semantics (built with -ffp-contract=off), one
// expression per reference expression in the same order, transcendentals from
// kpt_dmath.h.  Each routine cites the reference file:line it restates.
//
// Defined where the reference is not:
//  * Histogram sums are exact fixed point (weight * 2^23 rounded to an
//    integer, summed in 64 bits) instead of float atomicAdd in an unspecified
//    order (histogram.cu:84-89); every pixel of every block is counted (the
//    reference leaves LDS bins of partial 16x16 blocks uninitialised).
//  * The histogram average reads all 256 bins (the reference skips bins whose
//    thread falls outside the frame when the width is below 256).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "kpt_oracle.h"
#include "../kittlespt_amd/csrc/common/kpt_dmath.h"

namespace orp {

using kpt::dm::exp2f;
using kpt::dm::expf;
using kpt::dm::log2f;
using kpt::dm::powf;

struct F3 { float x, y, z; };
struct F4 { float x, y, z, w; };

static float Fmin(float a, float b) { return std::fmin(a, b); }
static float Fmax(float a, float b) { return std::fmax(a, b); }
// vector_types_extension.cuh:1168-1171
static float Clamp(float f, float a, float b) { return Fmax(a, Fmin(f, b)); }
// vector_types_extension.cuh:1146-1149 (a + t*(b - a)), per component
static float Lerp(float a, float b, float t) { return a + t * (b - a); }
static F4 Lerp4(F4 a, F4 b, float t) {
    return F4{Lerp(a.x, b.x, t), Lerp(a.y, b.y, t), Lerp(a.z, b.z, t), Lerp(a.w, b.w, t)};
}
static F3 Lerp3(F3 a, F3 b, float t) { return F3{Lerp(a.x, b.x, t), Lerp(a.y, b.y, t), Lerp(a.z, b.z, t)}; }
static float Dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static bool Bad(float v) { return std::isnan(v) || std::isinf(v); }

// clampOutput (vector_maths.cuh:27-34): NaN/Inf anywhere -> 0, else clamp [0, 1e8]
static F3 ClampOutput(F3 v) {
    if (Bad(v.x) || Bad(v.y) || Bad(v.z)) return F3{0.0f, 0.0f, 0.0f};
    return F3{Clamp(v.x, 0.0f, 1.0e8f), Clamp(v.y, 0.0f, 1.0e8f), Clamp(v.z, 0.0f, 1.0e8f)};
}

// ----------------------------------------------------------------- images
// DeviceTextureBuffer over a row-major RGBA32F image (device_texture_buffer.cuh:19-80)
struct Img {
    F4* p;
    int w, h;
    // textureReadNearest: truncate to int, clamp to the edge (cudaBoundaryModeClamp)
    F4 nearest(float x, float y) const {
        int ix = (int)x, iy = (int)y;
        ix = ix < 0 ? 0 : (ix > w - 1 ? w - 1 : ix);
        iy = iy < 0 ? 0 : (iy > h - 1 ? h - 1 : iy);
        return p[(size_t)iy * w + ix];
    }
    // textureReadBilinear (:37-71): taps at trunc(p) and +1, weights p - trunc(p)
    F4 bilinear(float x, float y, bool filter_alpha) const {
        int dx = (int)x, dy = (int)y;
        int s0 = dx, s1 = dx + 1, t0 = dy, t1 = dy + 1;
        float ws = x - (float)dx;
        float wt = y - (float)dy;
        F4 cp0 = nearest((float)s0, (float)t0);
        F4 cp1 = nearest((float)s1, (float)t0);
        F4 cp2 = nearest((float)s0, (float)t1);
        F4 cp3 = nearest((float)s1, (float)t1);
        F4 tc0 = Lerp4(cp0, cp1, ws);
        F4 tc1 = Lerp4(cp2, cp3, ws);
        F4 fc = Lerp4(tc0, tc1, wt);
        if (!filter_alpha) fc.w = cp0.w;
        return fc;
    }
};

// ----------------------------------------------------------------- colour
static const F3 kY709{0.2126f, 0.7152f, 0.0722f};   // sRGB_Y_coeffs, color.cuh:109
// RGBSpectrum::getLuminance (color.cuh:348-351)
static float Luminance(F3 c) { return (kY709.x * c.x) + (kY709.y * c.y) + (kY709.z * c.z); }

// convertRGB2XYZ / convertXYZ2RGB (color.cuh:33-52); literals rounded to float
static F3 RGB2XYZ(F3 c) {
    return F3{Dot3(F3{(float)0.4124564, (float)0.3575761, (float)0.1804375}, c),
              Dot3(F3{(float)0.2126729, (float)0.7151522, (float)0.0721750}, c),
              Dot3(F3{(float)0.0193339, (float)0.1191920, (float)0.9503041}, c)};
}
static F3 XYZ2RGB(F3 c) {
    return F3{Dot3(F3{(float)3.2404542, (float)-1.5371385, (float)-0.4985314}, c),
              Dot3(F3{(float)-0.9692660, (float)1.8760108, (float)0.0415560}, c),
              Dot3(F3{(float)0.0556434, (float)-0.2040259, (float)1.0572252}, c)};
}
// convertXYZ2Yxy (color.cuh:54-60)
static F3 XYZ2Yxy(F3 v) {
    float inv = 1.0f / Dot3(v, F3{1.0f, 1.0f, 1.0f});
    return F3{v.y, v.x * inv, v.y * inv};
}
// convertYxy2XYZ (color.cuh:62-72): z in double (1.0 literal)
static F3 Yxy2XYZ(F3 v) {
    F3 o;
    o.x = v.x * v.y / v.z;
    o.y = v.x;
    o.z = (float)((double)v.x * (1.0 - (double)v.y - (double)v.z) / (double)v.z);
    return o;
}

// Mat3 * float3 with the reference's column-major constructor
// (matrix_maths.cuh:40-47, 91-99): arguments are columns
struct M3 { F3 c0, c1, c2; };
static F3 Mul(const M3& m, F3 v) {
    return F3{m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z, m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
              m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z};
}
// film.cuh:35-51
static const M3 kSrgbToRec2020{{0.6274f, 0.0691f, 0.0164f}, {0.3293f, 0.9195f, 0.0880f}, {0.0433f, 0.0113f, 0.8956f}};
static const M3 kAgxInset{{0.856627153315983f, 0.137318972929847f, 0.11189821299995f},
                          {0.0951212405381588f, 0.761241990602591f, 0.0767994186031903f},
                          {0.0482516061458583f, 0.101439036467562f, 0.811302368396859f}};
static const M3 kAgxOutset{{1.9648846919172409596f, -0.29937618452442253746f, -0.16440106280678278299f},
                           {-0.85594737466675834968f, 1.3263980951083531115f, -0.23819967517076844919f},
                           {-0.10883731725048386702f, -0.02702191058393112346f, 1.4025007379775505276f}};
static const float kMiddleGray = 0.18f;

// AgXDefaultContrastApprox (film.cuh:61-72), per component, left to right
static float AgxCurve(float x) {
    float x2 = x * x;
    float x4 = x2 * x2;
    return +15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x -
           0.00232f;
}
// AgXFitted (film.cuh:74-116)
static F3 AgxFitted(F3 c, float white_ev, float black_ev) {
    const float min_ev = log2f(exp2f(black_ev) * kMiddleGray);
    const float max_ev = log2f(exp2f(white_ev) * kMiddleGray);
    const float range = max_ev - min_ev;
    F3 r = Mul(kSrgbToRec2020, c);
    r = F3{Fmax(r.x, 0.0f), Fmax(r.y, 0.0f), Fmax(r.z, 0.0f)};
    F3 ins = Mul(kAgxInset, r);
    F3 l = F3{Clamp(log2f(ins.x), min_ev, max_ev), Clamp(log2f(ins.y), min_ev, max_ev),
              Clamp(log2f(ins.z), min_ev, max_ev)};
    F3 n = F3{(l.x - min_ev) / range, (l.y - min_ev) / range, (l.z - min_ev) / range};
    return F3{AgxCurve(n.x), AgxCurve(n.y), AgxCurve(n.z)};
}
// AgXLook, Punchy (film.cuh:122-148): power 1.05, saturation 1.25
static F3 AgxLook(F3 v) {
    float luma = Dot3(v, kY709);
    const float power = (float)1.05, sat = 1.25f;
    // val * slope + offset with slope 1 and offset 0
    F3 s = F3{v.x * 1.0f + 0.0f, v.y * 1.0f + 0.0f, v.z * 1.0f + 0.0f};
    F3 nv = F3{powf(s.x, power), powf(s.y, power), powf(s.z, power)};
    return F3{luma + sat * (nv.x - luma), luma + sat * (nv.y - luma), luma + sat * (nv.z - luma)};
}
// sRGB OETF (color.cuh:419-424)
static float SrgbEncode(float v) {
    const float V = 0.0031308f, A = 12.92f, C = 0.055f, T = 2.4f;
    return (v <= V) ? (A * v) : ((1 + C) * powf(v, 1.0f / T) - C);
}
// AgXFittedEOTF (film.cuh:155-174)
static F3 AgxEotf(F3 v) {
    const float g = (float)2.4;
    F3 lin = F3{powf(v.x, g), powf(v.y, g), powf(v.z, g)};
    F3 o = Mul(kAgxOutset, lin);
    o = F3{Fmax(o.x, 0.0f), Fmax(o.y, 0.0f), Fmax(o.z, 0.0f)};
    return F3{SrgbEncode(o.x), SrgbEncode(o.y), SrgbEncode(o.z)};
}
// Film::computeNormalizedNonLinearSRGB (film.cuh:180-188)
static F3 Film(F3 c, float white_ev, float black_ev) {
    return AgxEotf(AgxLook(AgxFitted(c, white_ev, black_ev)));
}

// computePostProcess (kernels.cu:475-504) for one pixel
static F3 PostPixel(F3 c, const F3* bloom, float final_blend, float scalar, float white_ev, float black_ev) {
    if (bloom) c = Lerp3(c, *bloom, final_blend);
    F3 yxy = XYZ2Yxy(RGB2XYZ(c));
    yxy.x *= scalar;
    c = XYZ2RGB(Yxy2XYZ(yxy));
    return Film(c, white_ev, black_ev);
}

// ----------------------------------------------------------------- exposure
static const uint32_t kBins = 256;               // Constants::HISTOGRAM_SIZE
static const float kLumEps = 0.005f;             // Constants::HISTOGRAM_LUMINANCE_EPSILON
static const double kFix = 8388608.0;            // 2^23 fixed-point weight scale

// luminanceToBin (histogram.cu:18-33)
static uint32_t LumToBin(float lum, float min_log, float inv_range) {
    if (lum < kLumEps) return 0u;
    float n = Clamp((log2f(lum) - min_log) * inv_range, 0.0f, 1.0f);
    return (uint32_t)((n * (float)(kBins - 2)) + 1);
}
// centerMeteringWeight (histogram.cu:35-48); int2 / float truncates
static float MeterWeight(int W, int H, int px, int py, float radius_factor) {
    int cx = (int)((float)W / 2.0f), cy = (int)((float)H / 2.0f);
    float dx = (float)(cx - px), dy = (float)(cy - py);
    float distance = std::sqrt(dx * dx + dy * dy);
    distance *= (1.0f / radius_factor);
    float max_distance = Fmin((float)W, (float)H) / 2.0f;
    float nd = distance / max_distance;
    float y = Clamp((nd - 0.0f) / (1.0f - 0.0f), 0.0f, 1.0f);   // smoothstep(0, 1, nd)
    float weight = 1.0f - (y * y * (3.0f - (2.0f * y)));
    return Clamp(weight, 0.0f, 1.0f);
}
static uint64_t ToFixed(float w) { return (uint64_t)(uint32_t)(w * 8388608.0f + 0.5f); }
static float FromFixed(uint64_t q) { return (float)((double)q / kFix); }

// histogramComputeKernel (histogram.cu:51-90) over the modulated render
static void Histogram(const F4* img, int W, int H, float white_ev, float black_ev, uint64_t bins[256]) {
    std::memset(bins, 0, sizeof(uint64_t) * kBins);
    float range = white_ev - black_ev;
    float inv = 1.0f / range;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const F4& c = img[(size_t)y * W + x];
            uint32_t b = LumToBin(Luminance(F3{c.x, c.y, c.z}), black_ev, inv);
            float w = MeterWeight(W, H, x, y, 0.4f);
            w = (b > 0) ? w : 1.0f;
            bins[b] += ToFixed(w);
        }
}
// histogramAverageLuminanceComputeKernel (histogram.cu:93-153)
static float HistogramAverage(const uint64_t bins[256], float white_ev, float black_ev, float last, float dt_ms) {
    float nw[256], nwl[256];
    for (uint32_t i = 0; i < kBins; ++i) {
        float v = FromFixed(bins[i]);
        nw[i] = v;
        nwl[i] = v * (float)(int)i;
    }
    for (uint32_t cut = kBins / 2; cut > 0; cut >>= 1)
        for (uint32_t i = 0; i < cut; ++i) {
            nwl[i] += nwl[i + cut];
            nw[i] += nw[i + cut];
        }
    float black_count = FromFixed(bins[0]);
    float non_black = (float)std::fmax((double)(nw[0] - black_count), 0.001);
    float avg_bin = nwl[0] / non_black;
    float range = white_ev - black_ev;
    float avg_lum = powf(2.0f, (Clamp((avg_bin - 1.0f) / (float)(kBins - 2), 0.0f, 1.0f) * range) + black_ev);
    const float speed = 0.5f;
    return Lerp(last, avg_lum, 1.0f - expf(-dt_ms * speed));
}
// AutoExposureProgram::computeExposure (autoexposure_program.hpp:18-24, 60-90)
static void ExposureProgram(kpt_exposure* e, float avg_lum) {
    const float K = 12.5f;
    float ev_target = log2f((avg_lum * 100.0f) / K) - e->ev_comp;
    ev_target += 6.0f;
    const float focal = 0.1f;
    float aperture = e->aperture_f_num;
    float shutter = 1.0f / (focal * 1000.0f);
    float iso = Clamp((aperture * aperture * 100.0f) / (shutter * powf(2.0f, ev_target)), (float)e->min_iso,
                      (float)e->max_iso);
    float ev = log2f((aperture * aperture * 100.0f) / (shutter * iso));
    float ev_diff = ev_target - ev;
    shutter = Clamp(shutter * powf(2.0f, -ev_diff), e->min_shutter_secs, e->max_shutter_secs);
    e->shutter_speed_secs = shutter;
    e->iso = iso;
}
// getStandardOutputBasedExposure (autoexposure_program.hpp:38-47) + Renderer::setExposure
static void SetExposure(const kpt_exposure* e, kpt_film_state* f) {
    const float q = 0.65f;
    float l_avg = (10.0f / q) * ((e->aperture_f_num * e->aperture_f_num) / (e->iso * e->shutter_speed_secs));
    f->luminance_exposure_scalar = 0.18f / l_avg;
    f->white_point_ev = e->white_point_ev;
    f->black_point_ev = e->black_point_ev;
}

// ----------------------------------------------------------------- bloom
// karisAverage (bloom.cu:6-17)
static F4 Karis(F4 a, F4 b, F4 c, F4 d) {
    float w0 = 1.0f / Luminance(F3{a.x, a.y, a.z});
    float w1 = 1.0f / Luminance(F3{b.x, b.y, b.z});
    float w2 = 1.0f / Luminance(F3{c.x, c.y, c.z});
    float w3 = 1.0f / Luminance(F3{d.x, d.y, d.z});
    float net = 1.0f / (w0 + w1 + w2 + w3);
    auto mul = [](F4 v, float s) { return F4{v.x * s, v.y * s, v.z * s, v.w * s}; };
    auto add = [](F4 u, F4 v) { return F4{u.x + v.x, u.y + v.y, u.z + v.z, u.w + v.w}; };
    return mul(add(add(add(mul(a, w0), mul(b, w1)), mul(c, w2)), mul(d, w3)), net);
}
static F4 Box(F4 a, F4 b, F4 c, F4 d, bool karis) {
    return karis ? Karis(a, b, c, d) : Lerp4(Lerp4(a, b, 0.5f), Lerp4(c, d, 0.5f), 0.5f);
}
// textureRead36Texels (bloom.cu:19-84)
static F4 Read36(const Img& t, float px, float py, bool karis) {
    F4 a = t.bilinear(px + -2.0f, py + 2.0f, true);
    F4 b = t.bilinear(px + 0.0f, py + 2.0f, true);
    F4 c = t.bilinear(px + 2.0f, py + 2.0f, true);
    F4 d = t.bilinear(px + -2.0f, py + 0.0f, true);
    F4 e = t.bilinear(px, py, true);
    F4 f = t.bilinear(px + 2.0f, py + 0.0f, true);
    F4 g = t.bilinear(px + -2.0f, py + -2.0f, true);
    F4 h = t.bilinear(px + 0.0f, py + -2.0f, true);
    F4 i = t.bilinear(px + 2.0f, py + -2.0f, true);
    F4 hb1 = Box(a, b, d, e, karis);
    F4 hb2 = Box(b, c, e, f, karis);
    F4 hb3 = Box(d, e, g, h, karis);
    F4 hb4 = Box(e, f, h, i, karis);
    F4 j = t.bilinear(px + -1.0f, py + 1.0f, true);
    F4 k = t.bilinear(px + 1.0f, py + 1.0f, true);
    F4 l = t.bilinear(px + -1.0f, py + -1.0f, true);
    F4 m = t.bilinear(px + 1.0f, py + -1.0f, true);
    F4 hb5 = Box(j, k, l, m, karis);
    return Lerp4(hb5, Lerp4(Lerp4(hb1, hb2, 0.5f), Lerp4(hb3, hb4, 0.5f), 0.5f), 0.5f);
}
// downSample (bloom.cu:94-114)
static void DownSample(const Img& src, Img& dst, bool karis) {
    for (int y = 0; y < dst.h; ++y)
        for (int x = 0; x < dst.w; ++x) {
            float u = (float)x / (float)dst.w, v = (float)y / (float)dst.h;
            float sx = u * (float)src.w, sy = v * (float)src.h;
            F4 c = Read36(src, sx, sy, karis);
            F3 o = ClampOutput(F3{c.x, c.y, c.z});
            dst.p[(size_t)y * dst.w + x] = F4{o.x, o.y, o.z, 1.0f};
        }
}
// upSampleCombine (bloom.cu:116-145)
static void UpSample(const Img& src, Img& dst, float internal_blend) {
    static const float K[3][3] = {{1.0f / 16.0f, 2.0f / 16.0f, 1.0f / 16.0f},
                                  {2.0f / 16.0f, 4.0f / 16.0f, 2.0f / 16.0f},
                                  {1.0f / 16.0f, 2.0f / 16.0f, 1.0f / 16.0f}};
    for (int y = 0; y < dst.h; ++y)
        for (int x = 0; x < dst.w; ++x) {
            float u = (float)x / (float)dst.w, v = (float)y / (float)dst.h;
            float sx = u * (float)src.w, sy = v * (float)src.h;
            F4 acc{0.0f, 0.0f, 0.0f, 0.0f};
            for (int ty = -1; ty <= 1; ++ty)
                for (int tx = -1; tx <= 1; ++tx) {
                    F4 c = src.bilinear(sx + (float)tx, sy + (float)ty, true);
                    float k = K[ty + 1][tx + 1];
                    acc = F4{acc.x + k * c.x, acc.y + k * c.y, acc.z + k * c.z, acc.w + k * c.w};
                }
            F4 prev = dst.nearest((float)x, (float)y);
            acc = Lerp4(prev, acc, internal_blend);
            F3 o = ClampOutput(F3{acc.x, acc.y, acc.z});
            dst.p[(size_t)y * dst.w + x] = F4{o.x, o.y, o.z, 1.0f};
        }
}
// MipChain::resize / getMaxValidMipLevels (mipchain.cuh:17-50): levels beyond mip 0
static int MipLevels(int W, int H) {
    int mx = (int)std::log2((double)W), my = (int)std::log2((double)H);
    int l = mx < my ? mx : my;
    return l < 6 ? l : 6;
}
// Renderer::executeBloomGeneration (renderer.cu:438-475): returns mip 0
static void Bloom(const F4* base, int W, int H, bool karis, float internal_blend, std::vector<std::vector<F4>>& mips) {
    int L = MipLevels(W, H);
    mips.assign(L + 1, {});
    for (int k = 0; k <= L; ++k) mips[k].assign((size_t)(W >> k) * (H >> k), F4{0, 0, 0, 0});
    std::memcpy(mips[0].data(), base, sizeof(F4) * (size_t)W * H);
    for (int k = 0; k < L; ++k) {
        Img s{mips[k].data(), W >> k, H >> k}, d{mips[k + 1].data(), W >> (k + 1), H >> (k + 1)};
        DownSample(s, d, karis && k == 0);
    }
    for (int k = L; k > 0; --k) {
        Img s{mips[k].data(), W >> k, H >> k}, d{mips[k - 1].data(), W >> (k - 1), H >> (k - 1)};
        UpSample(s, d, internal_blend);
    }
}

}  // namespace orp

using namespace orp;

extern "C" {

// ----------------------------------------------------------------- FSR 1.0
// The reference's USE_FSR build of the upscale pass (upscale.cuh:13-258 EASU,
// 260-298 upscale, 302-378 RCAS, 389-408 sharpen; kernels.cu:84-103 runs EASU
// render -> output, RCAS output -> back buffer; renderer.cu:138-147 copies the
// back buffer to the output), restated expression by expression: float
// vector arithmetic as vector_types_extension.cuh defines it, the double
// literals' promotions kept, inversesqrt as the IEEE 1/sqrtf (the reference's
// rsqrtf is an approximation), RCAS without FSR_RCAS_DENOISE and sharpness 0
// (con = exp2(-0) = 1).
struct F2 {
    float x, y;
};
static F3 EasuCF(F2 p, const Img& t) {   // textureReadBilinearUV(p, true): p * dimensions
    const F4 c = t.bilinear(p.x * (float)t.w, p.y * (float)t.h, true);
    return F3{c.x, c.y, c.z};
}
static float EasuLuma(F3 c) { return (float)((double)c.y + 0.5 * (double)(c.x + c.z)); }
static float HmClamp(float f, float a, float b) { return std::fmax(a, std::fmin(f, b)); }   // helper clamp
// FsrEasuSetF (upscale.cuh:103-130)
static void EasuSet(F2& dir, float& len, float w, float lA, float lB, float lC, float lD, float lE) {
    float lenX = std::fmax(std::fabs(lD - lC), std::fabs(lC - lB));
    const float dirX = lD - lB;
    dir.x += dirX * w;
    lenX = HmClamp(std::fabs(dirX) / lenX, 0.0f, 1.0f);
    lenX *= lenX;
    len += lenX * w;
    float lenY = std::fmax(std::fabs(lE - lC), std::fabs(lC - lA));
    const float dirY = lE - lA;
    dir.y += dirY * w;
    lenY = HmClamp(std::fabs(dirY) / lenY, 0.0f, 1.0f);
    lenY *= lenY;
    len += lenY * w;
}
// FsrEasuTapF (upscale.cuh:64-101)
static void EasuTap(F3& aC, float& aW, F2 off, F2 dir, F2 len, float lob, float clp, F3 c) {
    F2 v{off.x * dir.x + off.y * dir.y, off.x * -dir.y + off.y * dir.x};
    v.x *= len.x;
    v.y *= len.y;
    const float d2 = std::fmin(v.x * v.x + v.y * v.y, clp);
    float wB = (float)(0.4 * (double)d2 - 1.0);
    float wA = (float)((double)(lob * d2) - 1.0);
    wB *= wB;
    wA *= wA;
    wB = (float)(1.5625 * (double)wB - 0.5625);
    const float w = wB * wA;
    aC = F3{aC.x + c.x * w, aC.y + c.y * w, aC.z + c.z * w};
    aW += w;
}
// FsrEasuCon + FsrEasuF for output pixel (x, y) of an OW x OH upscale of t.
static F3 EasuPixel(const Img& t, int OW, int OH, int x, int y) {
    const float iw = (float)t.w, ih = (float)t.h, ow = (float)OW, oh = (float)OH;
    const float c0x = iw / ow, c0y = ih / oh;
    const float c0z = (float)(0.5 * (double)iw / (double)ow - 0.5), c0w = (float)(0.5 * (double)ih / (double)oh - 0.5);
    const float c1x = 1.0f / iw, c1y = 1.0f / ih, c1z = 1.0f / iw, c1w = -1.0f / ih;
    const float c2x = -1.0f / iw, c2y = 2.0f / ih, c2z = 1.0f / iw, c2w = 2.0f / ih;
    const float c3x = 0.0f / iw, c3y = 4.0f / ih;
    F2 pp{(float)x * c0x + c0z, (float)y * c0y + c0w};
    const F2 fp{std::floor(pp.x), std::floor(pp.y)};
    pp = F2{pp.x - fp.x, pp.y - fp.y};
    const F2 p0{fp.x * c1x + c1z, fp.y * c1y + c1w};
    const F2 p1{p0.x + c2x, p0.y + c2y}, p2{p0.x + c2z, p0.y + c2w}, p3{p0.x + c3x, p0.y + c3y};
    const float ox = -0.5f * c1x, oy = 0.5f * c1x, oz = -0.5f * c1y, ow_ = 0.5f * c1y;
    const F3 bC = EasuCF(F2{p0.x + ox, p0.y + ow_}, t), cC = EasuCF(F2{p0.x + oy, p0.y + ow_}, t);
    const F3 iC = EasuCF(F2{p1.x + ox, p1.y + ow_}, t), jC = EasuCF(F2{p1.x + oy, p1.y + ow_}, t);
    const F3 fC = EasuCF(F2{p1.x + oy, p1.y + oz}, t), eC = EasuCF(F2{p1.x + ox, p1.y + oz}, t);
    const F3 kC = EasuCF(F2{p2.x + ox, p2.y + ow_}, t), lC = EasuCF(F2{p2.x + oy, p2.y + ow_}, t);
    const F3 hC = EasuCF(F2{p2.x + oy, p2.y + oz}, t), gC = EasuCF(F2{p2.x + ox, p2.y + oz}, t);
    const F3 oC = EasuCF(F2{p3.x + oy, p3.y + oz}, t), nC = EasuCF(F2{p3.x + ox, p3.y + oz}, t);
    const float bL = EasuLuma(bC), cL = EasuLuma(cC), iL = EasuLuma(iC), jL = EasuLuma(jC), fL = EasuLuma(fC),
                eL = EasuLuma(eC), kL = EasuLuma(kC), lL = EasuLuma(lC), hL = EasuLuma(hC), gL = EasuLuma(gC),
                oL = EasuLuma(oC), nL = EasuLuma(nC);
    F2 dir{0.0f, 0.0f};
    float len = 0.0f;
    EasuSet(dir, len, (float)((1.0 - (double)pp.x) * (1.0 - (double)pp.y)), bL, eL, fL, gL, jL);
    EasuSet(dir, len, (float)((double)pp.x * (1.0 - (double)pp.y)), cL, fL, gL, hL, kL);
    EasuSet(dir, len, (float)((1.0 - (double)pp.x) * (double)pp.y), fL, iL, jL, kL, nL);
    EasuSet(dir, len, pp.x * pp.y, gL, jL, kL, lL, oL);
    const F2 dir2{dir.x * dir.x, dir.y * dir.y};
    float dirR = dir2.x + dir2.y;
    const bool zro = (double)dirR < (1.0 / 32768.0);
    dirR = 1.0f / std::sqrt(dirR);
    dirR = zro ? 1.0f : dirR;
    dir.x = zro ? 1.0f : dir.x;
    dir = F2{dir.x * dirR, dir.y * dirR};
    len = (float)((double)len * 0.5);
    len *= len;
    const float stretch = (dir.x * dir.x + dir.y * dir.y) / std::fmax(std::fabs(dir.x), std::fabs(dir.y));
    const F2 len2{(float)(1.0 + ((double)stretch - 1.0) * (double)len), (float)(1.0 - 0.5 * (double)len)};
    const float lob = (float)(0.5 - 0.29 * (double)len);
    const float clp = (float)(1.0 / (double)lob);
    const F3 min4{std::fmin(std::fmin(fC.x, gC.x), std::fmin(jC.x, kC.x)),
                  std::fmin(std::fmin(fC.y, gC.y), std::fmin(jC.y, kC.y)),
                  std::fmin(std::fmin(fC.z, gC.z), std::fmin(jC.z, kC.z))};
    const F3 max4{std::fmax(std::fmax(fC.x, gC.x), std::fmax(jC.x, kC.x)),
                  std::fmax(std::fmax(fC.y, gC.y), std::fmax(jC.y, kC.y)),
                  std::fmax(std::fmax(fC.z, gC.z), std::fmax(jC.z, kC.z))};
    F3 aC{0.0f, 0.0f, 0.0f};
    float aW = 0.0f;
    EasuTap(aC, aW, F2{0.0f - pp.x, -1.0f - pp.y}, dir, len2, lob, clp, bC);
    EasuTap(aC, aW, F2{1.0f - pp.x, -1.0f - pp.y}, dir, len2, lob, clp, cC);
    EasuTap(aC, aW, F2{-1.0f - pp.x, 1.0f - pp.y}, dir, len2, lob, clp, iC);
    EasuTap(aC, aW, F2{0.0f - pp.x, 1.0f - pp.y}, dir, len2, lob, clp, jC);
    EasuTap(aC, aW, F2{0.0f - pp.x, 0.0f - pp.y}, dir, len2, lob, clp, fC);
    EasuTap(aC, aW, F2{-1.0f - pp.x, 0.0f - pp.y}, dir, len2, lob, clp, eC);
    EasuTap(aC, aW, F2{1.0f - pp.x, 1.0f - pp.y}, dir, len2, lob, clp, kC);
    EasuTap(aC, aW, F2{2.0f - pp.x, 1.0f - pp.y}, dir, len2, lob, clp, lC);
    EasuTap(aC, aW, F2{2.0f - pp.x, 0.0f - pp.y}, dir, len2, lob, clp, hC);
    EasuTap(aC, aW, F2{1.0f - pp.x, 0.0f - pp.y}, dir, len2, lob, clp, gC);
    EasuTap(aC, aW, F2{1.0f - pp.x, 2.0f - pp.y}, dir, len2, lob, clp, oC);
    EasuTap(aC, aW, F2{0.0f - pp.x, 2.0f - pp.y}, dir, len2, lob, clp, nC);
    const F3 r{aC.x / aW, aC.y / aW, aC.z / aW};
    return F3{std::fmin(max4.x, std::fmax(min4.x, r.x)), std::fmin(max4.y, std::fmax(min4.y, r.y)),
              std::fmin(max4.z, std::fmax(min4.z, r.z))};
}
// FsrRcasF (upscale.cuh:316-378) at output pixel (x, y) of t, con = 1.
static F3 RcasPixel(const Img& t, int x, int y) {
    const float w = (float)t.w, h = (float)t.h;
    auto load = [&](float px, float py) {   // textureReadBilinearUV(p / dimensions, true)
        const float u = px / w, v = py / h;
        const F4 c = t.bilinear(u * w, v * h, true);
        return F3{c.x, c.y, c.z};
    };
    const float sx = (float)x, sy = (float)y;
    const F3 b = load(sx + 0.0f, sy + -1.0f), d = load(sx + -1.0f, sy + 0.0f), e = load(sx, sy),
             f = load(sx + 1.0f, sy + 0.0f), hh = load(sx + 0.0f, sy + 1.0f);
    const F3 mn4{std::fmin(b.x, std::fmin(f.x, hh.x)), std::fmin(b.y, std::fmin(f.y, hh.y)),
                 std::fmin(b.z, std::fmin(f.z, hh.z))};
    const F3 mx4{std::fmax(b.x, std::fmax(f.x, hh.x)), std::fmax(b.y, std::fmax(f.y, hh.y)),
                 std::fmax(b.z, std::fmax(f.z, hh.z))};
    const F3 hitMin{mn4.x / (4.0f * mx4.x), mn4.y / (4.0f * mx4.y), mn4.z / (4.0f * mx4.z)};
    const F3 hitMax{(1.0f - mx4.x) / (4.0f * mn4.x + -4.0f), (1.0f - mx4.y) / (4.0f * mn4.y + -4.0f),
                    (1.0f - mx4.z) / (4.0f * mn4.z + -4.0f)};
    const F3 lobeRGB{std::fmax(-hitMin.x, hitMax.x), std::fmax(-hitMin.y, hitMax.y), std::fmax(-hitMin.z, hitMax.z)};
    const float m = std::fmax(lobeRGB.x, std::fmax(lobeRGB.y, lobeRGB.z));
    const float lobe = (float)(std::fmax(-(0.25 - (1.0 / 16.0)), std::fmin((double)m, 0.0)) * 1.0);
    const F3 s{((b.x + d.x) + hh.x) + f.x, ((b.y + d.y) + hh.y) + f.y, ((b.z + d.z) + hh.z) + f.z};
    const float den = (float)(4.0 * (double)lobe + 1.0);
    return F3{(lobe * s.x + e.x) / den, (lobe * s.y + e.y) / den, (lobe * s.z + e.z) / den};
}
// EASU of src (W x H) to OW x OH, then RCAS of that; out OW x OH RGBA (alpha 1).
static void FsrUpscale(const Img& src, int OW, int OH, std::vector<F4>& out) {
    std::vector<F4> easu((size_t)OW * OH);
    for (int y = 0; y < OH; ++y)
        for (int x = 0; x < OW; ++x) {
            const F3 c = EasuPixel(src, OW, OH, x, y);
            easu[(size_t)y * OW + x] = F4{c.x, c.y, c.z, 1.0f};
        }
    const Img e{easu.data(), OW, OH};
    out.resize((size_t)OW * OH);
    for (int y = 0; y < OH; ++y)
        for (int x = 0; x < OW; ++x) {
            const F3 c = RcasPixel(e, x, y);
            out[(size_t)y * OW + x] = F4{c.x, c.y, c.z, 1.0f};
        }
}

// executeRendering's post chain with the render size W x H and the output
// size OW x OH: modulation and auto exposure at the render size; the output
// texture is the render buffer (same size, no upscale) or its bilinear upscale
// (upscale.cuh:287-292: output pixel (x, y) reads textureReadBilinear at
// (x / OW * W, y / OH * H), alpha filtered); bloom and the post pass at the
// output size.  out_output and out_bloom hold OW x OH pixels.
void orc_post_frame_sized(const float* estimate, const uint32_t* gbuffer, int32_t W, int32_t H, int32_t OW,
                          int32_t OH, int32_t upscale, const kpt_settings* s, float frame_delta_ms,
                          kpt_exposure* exposure, kpt_film_state* film, float* out_render, float* out_output,
                          float* out_bloom, uint64_t* out_hist) {
    const size_t n = (size_t)W * H;
    std::vector<F4> render(n);
    // modulateSamples (kernels.cu:453-472): render *= 8-bit albedo (byteToFloat, packing.cuh:19-22)
    for (size_t i = 0; i < n; ++i) {
        uint32_t a = gbuffer[4 * i + 2];
        F3 alb{(float)(a & 0xFFu) / 255.0f, (float)((a >> 8) & 0xFFu) / 255.0f, (float)((a >> 16) & 0xFFu) / 255.0f};
        const float* e = estimate + 4 * i;
        render[i] = F4{e[0] * alb.x, e[1] * alb.y, e[2] * alb.z, 1.0f};
    }
    // auto exposure (renderer.cu:125-132)
    if (s->enable_auto_exposure) {
        uint64_t bins[256];
        Histogram(render.data(), W, H, film->white_point_ev, film->black_point_ev, bins);
        if (out_hist) std::memcpy(out_hist, bins, sizeof(bins));
        film->scene_average_luminance = HistogramAverage(bins, film->white_point_ev, film->black_point_ev,
                                                         film->scene_average_luminance, frame_delta_ms);
        ExposureProgram(exposure, film->scene_average_luminance);
        SetExposure(exposure, film);
    }
    if (out_render) std::memcpy(out_render, render.data(), n * 16);
    // output <- render, or its upscale (renderer.cu:138-156), bloom (:159-162), post (:164)
    const bool up = upscale || OW != W || OH != H;
    const size_t no = (size_t)OW * OH;
    std::vector<F4> outtex(no);
    if (up && s->upscale_fsr) {
        Img src{render.data(), W, H};
        FsrUpscale(src, OW, OH, outtex);
    } else if (up) {
        Img src{render.data(), W, H};
        for (int y = 0; y < OH; ++y)
            for (int x = 0; x < OW; ++x) {
                const float u = (float)x / (float)OW, v = (float)y / (float)OH;
                outtex[(size_t)y * OW + x] = src.bilinear(u * (float)W, v * (float)H, true);
            }
    } else {
        outtex = render;
    }
    std::vector<std::vector<F4>> mips;
    const bool bloom = s->bloom_generate_bloom != 0;
    if (bloom) Bloom(outtex.data(), OW, OH, s->bloom_use_karis_average != 0, s->bloom_internal_blend, mips);
    for (size_t i = 0; i < no; ++i) {
        F3 c{outtex[i].x, outtex[i].y, outtex[i].z};
        F3 b{0, 0, 0};
        if (bloom) b = F3{mips[0][i].x, mips[0][i].y, mips[0][i].z};
        F3 o = PostPixel(c, bloom ? &b : nullptr, s->bloom_final_blend, film->luminance_exposure_scalar,
                         film->white_point_ev, film->black_point_ev);
        float ov[4] = {o.x, o.y, o.z, 1.0f};
        if (out_output) std::memcpy(out_output + 4 * i, ov, 16);
        if (out_bloom && bloom) std::memcpy(out_bloom + 4 * i, &mips[0][i], 16);
    }
}

void orc_post_frame(const float* estimate, const uint32_t* gbuffer, int32_t W, int32_t H, const kpt_settings* s,
                    float frame_delta_ms, kpt_exposure* exposure, kpt_film_state* film, float* out_render,
                    float* out_output, float* out_bloom, uint64_t* out_hist) {
    orc_post_frame_sized(estimate, gbuffer, W, H, W, H, 0, s, frame_delta_ms, exposure, film, out_render, out_output,
                         out_bloom, out_hist);
}

// FSR 1.0 alone (EASU then RCAS, the USE_FSR build) of a W x H RGBA image to
// OW x OH; out_easu (optional) receives the EASU pass before sharpening.
void orc_upscale_fsr(const float* src, int32_t W, int32_t H, int32_t OW, int32_t OH, float* out, float* out_easu) {
    std::vector<F4> s4((size_t)W * H);
    std::memcpy(s4.data(), src, s4.size() * 16);
    Img img{s4.data(), W, H};
    std::vector<F4> o;
    FsrUpscale(img, OW, OH, o);
    std::memcpy(out, o.data(), o.size() * 16);
    if (out_easu)
        for (int y = 0; y < OH; ++y)
            for (int x = 0; x < OW; ++x) {
                const F3 c = EasuPixel(img, OW, OH, x, y);
                const float v[4] = {c.x, c.y, c.z, 1.0f};
                std::memcpy(out_easu + 4 * ((size_t)y * OW + x), v, 16);
            }
}

// The RCAS pass alone (upscale.cuh:316-378, con = 1) of a W x H RGBA image.
void orc_fsr_rcas(const float* src, int32_t W, int32_t H, float* out) {
    std::vector<F4> s4((size_t)W * H);
    std::memcpy(s4.data(), src, s4.size() * 16);
    Img img{s4.data(), W, H};
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const F3 c = RcasPixel(img, x, y);
            const float v[4] = {c.x, c.y, c.z, 1.0f};
            std::memcpy(out + 4 * ((size_t)y * W + x), v, 16);
        }
}

// The upscale alone (upscale.cuh:287-292) of a W x H RGBA image to OW x OH.
void orc_upscale_bilinear(const float* src, int32_t W, int32_t H, int32_t OW, int32_t OH, float* out) {
    std::vector<F4> s4((size_t)W * H);
    std::memcpy(s4.data(), src, s4.size() * 16);
    Img img{s4.data(), W, H};
    for (int y = 0; y < OH; ++y)
        for (int x = 0; x < OW; ++x) {
            const float u = (float)x / (float)OW, v = (float)y / (float)OH;
            const F4 c = img.bilinear(u * (float)W, v * (float)H, true);
            std::memcpy(out + 4 * ((size_t)y * OW + x), &c, 16);
        }
}

void orc_set_exposure(const kpt_exposure* e, kpt_film_state* f) { SetExposure(e, f); }

void orc_exposure_program(kpt_exposure* e, float avg_lum) { ExposureProgram(e, avg_lum); }

void orc_lum_to_bin(const float* lum, int64_t n, float black_ev, float white_ev, uint32_t* out) {
    float inv = 1.0f / (white_ev - black_ev);
    for (int64_t i = 0; i < n; ++i) out[i] = LumToBin(lum[i], black_ev, inv);
}

void orc_meter_weight(int32_t W, int32_t H, const int32_t* xy, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = MeterWeight(W, H, xy[2 * i], xy[2 * i + 1], 0.4f);
}

void orc_histogram(const float* rgba, int32_t W, int32_t H, float white_ev, float black_ev, uint64_t* bins) {
    Histogram(reinterpret_cast<const F4*>(rgba), W, H, white_ev, black_ev, bins);
}

float orc_histogram_average(const uint64_t* bins, float white_ev, float black_ev, float last, float dt_ms) {
    return HistogramAverage(bins, white_ev, black_ev, last, dt_ms);
}

// computeEffects (kernels.cu:508-528) over a back buffer of W x H RGBA32F
// pixels (row 0 first), with getFishEyeUV (kernels.cu:340-376, power 1.65):
// fisheye coordinates, then R, G and B point-sampled at the coordinates plus,
// at, and minus the radial offset powf(|uv - 0.5| * 2, 2) * 0.01.  Reads
// follow textureReadNearestUV (device_texture_buffer.cuh:35-40): the float
// coordinate times the size, converted to int as the device converts (toward
// zero, saturating, NaN to 0), clamped to the edge; the write goes to
// make_int2(uv * size) of the pixel's own uv, dropped when out of range.
// Operation order as written there: 0.5 / prop in double narrowed to float,
// normalize as v * (1 / sqrt(dot)) (rsqrtf, IEEE), tan from kpt_dmath.h.
void orc_post_effects(const float* back, int32_t W, int32_t H, float* out) {
    auto f2i = [](float x) -> int {
        if (std::isnan(x)) return 0;
        if (x >= 2147483648.0f) return 2147483647;
        if (x <= -2147483648.0f) return (-2147483647 - 1);
        return (int)x;
    };
    const float fw = (float)W, fh = (float)H;
    auto tap = [&](float tu, float tv, int ch) {
        int ix = f2i(tu * fw), iy = f2i(tv * fh);
        ix = ix < 0 ? 0 : (ix > W - 1 ? W - 1 : ix);
        iy = iy < 0 ? 0 : (iy > H - 1 ? H - 1 : iy);
        return back[4 * ((size_t)iy * W + ix) + ch];
    };
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            // getFishEyeUV
            const float p_x = (float)x / fw, p_y = (float)y / fw;
            const float prop = fw / fh;
            const float m_x = 0.5f, m_y = (float)(0.5 / (double)prop);
            const float d_x = p_x - m_x, d_y = p_y - m_y;
            const float r = std::sqrt((d_x * d_x) + (d_y * d_y));
            const float power = 1.65f;
            const float bind = std::sqrt((m_x * m_x) + (m_y * m_y));
            const float inv = 1.0f / std::sqrt((d_x * d_x) + (d_y * d_y));
            const float n_x = d_x * inv, n_y = d_y * inv;
            const float t = kpt::dm::tanf(r * power), tb = kpt::dm::tanf(bind * power);
            const float u = m_x + (((n_x * t) * bind) / tb);
            const float v = (m_y + (((n_y * t) * bind) / tb)) * prop;
            // chromatic aberration
            const float uc = ((float)x + 0.5f) / fw, vc = ((float)y + 0.5f) / fh;
            const float o_x = kpt::dm::powf(std::fabs(uc - 0.5f) * 2.0f, 2.0f) * 0.01f;
            const float o_y = kpt::dm::powf(std::fabs(vc - 0.5f) * 2.0f, 2.0f) * 0.01f;
            const float rr = tap(u + o_x, v + o_y, 0);
            const float gg = tap(u, v, 1);
            const float bb = tap(u - o_x, v - o_y, 2);
            const int wx = f2i(uc * fw), wy = f2i(vc * fh);
            if (wx < 0 || wx >= W || wy < 0 || wy >= H) continue;
            float* o = out + 4 * ((size_t)wy * W + wx);
            o[0] = rr; o[1] = gg; o[2] = bb; o[3] = 1.0f;
        }
    }
}

void orc_film(const float* rgb, int64_t n, float scalar, float white_ev, float black_ev, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        F3 o = PostPixel(F3{rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]}, nullptr, 0.0f, scalar, white_ev, black_ev);
        out[3 * i] = o.x;
        out[3 * i + 1] = o.y;
        out[3 * i + 2] = o.z;
    }
}

void orc_agx_curve(const float* x, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = AgxCurve(x[i]);
}

void orc_srgb_encode(const float* x, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = SrgbEncode(x[i]);
}

void orc_yxy_roundtrip(const float* rgb, int64_t n, float scalar, float* out_yxy, float* out_rgb) {
    for (int64_t i = 0; i < n; ++i) {
        F3 y = XYZ2Yxy(RGB2XYZ(F3{rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]}));
        out_yxy[3 * i] = y.x;
        out_yxy[3 * i + 1] = y.y;
        out_yxy[3 * i + 2] = y.z;
        y.x *= scalar;
        F3 c = XYZ2RGB(Yxy2XYZ(y));
        out_rgb[3 * i] = c.x;
        out_rgb[3 * i + 1] = c.y;
        out_rgb[3 * i + 2] = c.z;
    }
}

void orc_bilinear(const float* rgba, int32_t W, int32_t H, const float* xy, int64_t n, int32_t filter_alpha,
                  float* out) {
    Img t{const_cast<F4*>(reinterpret_cast<const F4*>(rgba)), W, H};
    for (int64_t i = 0; i < n; ++i) {
        F4 c = t.bilinear(xy[2 * i], xy[2 * i + 1], filter_alpha != 0);
        std::memcpy(out + 4 * i, &c, 16);
    }
}

void orc_bloom_down(const float* src, int32_t sw, int32_t sh, float* dst, int32_t dw, int32_t dh, int32_t karis) {
    Img s{const_cast<F4*>(reinterpret_cast<const F4*>(src)), sw, sh};
    Img d{reinterpret_cast<F4*>(dst), dw, dh};
    DownSample(s, d, karis != 0);
}

void orc_bloom_up(const float* src, int32_t sw, int32_t sh, float* dst, int32_t dw, int32_t dh, float blend) {
    Img s{const_cast<F4*>(reinterpret_cast<const F4*>(src)), sw, sh};
    Img d{reinterpret_cast<F4*>(dst), dw, dh};
    UpSample(s, d, blend);
}

int32_t orc_bloom_levels(int32_t W, int32_t H) { return MipLevels(W, H); }

}  // extern "C"
