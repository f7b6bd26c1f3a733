// kpt_post.hip — the frame's post chain on gfx950: albedo modulation, the
// log-luminance histogram and auto exposure, the bloom mip chain, and the
// exposure + AgX film transform.
//
// Reference: Renderer::executeRendering after the path trace
// (renderer.cu:117-170) with modulateSamples (kernels.cu:453-472),
// histogramComputeKernel / histogramAverageLuminanceComputeKernel
// (histogram.cu:51-153), AutoExposureProgram (autoexposure_program.hpp:18-94)
// and Renderer::setExposure (renderer.cu:312-329), downSample /
// upSampleCombine (bloom.cu:94-145), computePostProcess (kernels.cu:475-504),
// Film / AgX (film.cuh:35-188), colour conversions (color.cuh:33-82, 409-424).
//
// MI355X layout of the work:
//  * one streaming pass (k_post_prepare) reads render + gbuffer once and does
//    the temporal copy, the modulation, the mip-0 copy and the histogram;
//    histogram weights are summed as 2^-23 fixed point in 64-bit LDS bins per
//    workgroup and then in 256 global bins, so the result does not depend on
//    the order of the atomics;
//  * the histogram average and the exposure program run in one 256-lane
//    workgroup and leave the exposure in device memory (no host round trip);
//  * bloom is 2 x levels small launches over mips at W>>k x H>>k;
//  * k_post_final reads render + mip 0 once and writes the output.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../common/kpt_dmath.h"
#include "kpt_launch.h"

namespace kpt {
namespace {

constexpr int POST_BLOCK = 256;
constexpr uint32_t HIST_BINS = 256;                 // Constants::HISTOGRAM_SIZE
constexpr float HIST_LUM_EPS = 0.005f;              // Constants::HISTOGRAM_LUMINANCE_EPSILON

struct c3 { float x, y, z; };
struct c4 { float x, y, z, w; };

__device__ __forceinline__ c3 C3(float x, float y, float z) { c3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ c4 C4(float x, float y, float z, float w) {
    c4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r;
}
__device__ __forceinline__ c4 ld(const float4& v) { return C4(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ float4 st(c4 v) { return make_float4(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ float fmin_(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ float fmax_(float a, float b) { return __builtin_fmaxf(a, b); }
// clamp (vector_types_extension.cuh:1168-1171)
__device__ __forceinline__ float clampf_(float f, float a, float b) { return fmax_(a, fmin_(f, b)); }
// lerp (vector_types_extension.cuh:1146-1161): a + t * (b - a)
__device__ __forceinline__ float lerpf_(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ c4 lerp4(c4 a, c4 b, float t) {
    return C4(lerpf_(a.x, b.x, t), lerpf_(a.y, b.y, t), lerpf_(a.z, b.z, t), lerpf_(a.w, b.w, t));
}
__device__ __forceinline__ float dot3(c3 a, c3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ bool bad(float v) { return __builtin_isnan(v) || __builtin_isinf(v); }
// clampOutput (vector_maths.cuh:27-34)
__device__ __forceinline__ c3 clamp_output3(c3 v) {
    if (bad(v.x) || bad(v.y) || bad(v.z)) return C3(0.0f, 0.0f, 0.0f);
    return C3(clampf_(v.x, 0.0f, 1.0e8f), clampf_(v.y, 0.0f, 1.0e8f), clampf_(v.z, 0.0f, 1.0e8f));
}
// RGBSpectrum::getLuminance (color.cuh:348-351)
__device__ __forceinline__ float luminance(float r, float g, float b) {
    return (0.2126f * r) + (0.7152f * g) + (0.0722f * b);
}

// ------------------------------------------------------------ texture reads
// DeviceTextureBuffer over a row-major RGBA32F image (device_texture_buffer.cuh:19-80)
struct Tex {
    const float4* p;
    int w, h;
    __device__ __forceinline__ c4 nearest_i(int ix, int iy) const {
        ix = ix < 0 ? 0 : (ix > w - 1 ? w - 1 : ix);
        iy = iy < 0 ? 0 : (iy > h - 1 ? h - 1 : iy);
        return ld(p[(size_t)iy * w + ix]);
    }
    // textureReadBilinear with filter_alpha: taps at trunc(p) and +1
    __device__ __forceinline__ c4 bilinear(float x, float y) const {
        const int dx = (int)x, dy = (int)y;
        const float ws = x - (float)dx, wt = y - (float)dy;
        c4 cp0 = nearest_i(dx, dy), cp1 = nearest_i(dx + 1, dy);
        c4 cp2 = nearest_i(dx, dy + 1), cp3 = nearest_i(dx + 1, dy + 1);
        return lerp4(lerp4(cp0, cp1, ws), lerp4(cp2, cp3, ws), wt);
    }
};

// ------------------------------------------------------------ colour / film
// matrix entries are double literals narrowed to float, as make_float3(double...) does
#define F_(x) ((float)(x))
__device__ __forceinline__ c3 rgb_to_xyz(c3 c) {
    return C3(dot3(C3(F_(0.4124564), F_(0.3575761), F_(0.1804375)), c),
              dot3(C3(F_(0.2126729), F_(0.7151522), F_(0.0721750)), c),
              dot3(C3(F_(0.0193339), F_(0.1191920), F_(0.9503041)), c));
}
__device__ __forceinline__ c3 xyz_to_rgb(c3 c) {
    return C3(dot3(C3(F_(3.2404542), F_(-1.5371385), F_(-0.4985314)), c),
              dot3(C3(F_(-0.9692660), F_(1.8760108), F_(0.0415560)), c),
              dot3(C3(F_(0.0556434), F_(-0.2040259), F_(1.0572252)), c));
}
__device__ __forceinline__ c3 xyz_to_Yxy(c3 v) {
    const float inv = 1.0f / dot3(v, C3(1.0f, 1.0f, 1.0f));
    return C3(v.y, v.x * inv, v.y * inv);
}
// convertYxy2XYZ: the z term evaluates in double (double literal 1.0)
__device__ __forceinline__ c3 Yxy_to_xyz(c3 v) {
    return C3(v.x * v.y / v.z, v.x, (float)((double)v.x * (1.0 - (double)v.y - (double)v.z) / (double)v.z));
}
// Mat3(c1r1..c3r3) * v with columns given first (matrix_maths.cuh:40-47, 91-99)
__device__ __forceinline__ c3 mat3_cols(const float* m, c3 v) {
    return C3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
              m[2] * v.x + m[5] * v.y + m[8] * v.z);
}
__constant__ float kSrgbTo2020[9] = {0.6274f, 0.0691f, 0.0164f, 0.3293f, 0.9195f, 0.0880f, 0.0433f, 0.0113f, 0.8956f};
__constant__ float kInset[9] = {0.856627153315983f, 0.137318972929847f, 0.11189821299995f,
                                0.0951212405381588f, 0.761241990602591f, 0.0767994186031903f,
                                0.0482516061458583f, 0.101439036467562f, 0.811302368396859f};
__constant__ float kOutset[9] = {1.9648846919172409596f, -0.29937618452442253746f, -0.16440106280678278299f,
                                 -0.85594737466675834968f, 1.3263980951083531115f, -0.23819967517076844919f,
                                 -0.10883731725048386702f, -0.02702191058393112346f, 1.4025007379775505276f};

// AgXDefaultContrastApprox (film.cuh:61-72)
__device__ __forceinline__ float agx_curve(float x) {
    const float x2 = x * x;
    const float x4 = x2 * x2;
    return +15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x -
           0.00232f;
}
// IEC 61966-2-1 OETF (color.cuh:419-424)
__device__ __forceinline__ float srgb_oetf(float v) {
    const float V = 0.0031308f, A = 12.92f, Cc = 0.055f, T = 2.4f;
    return (v <= V) ? (A * v) : ((1 + Cc) * dm::powf(v, 1.0f / T) - Cc);
}
// Film::computeNormalizedNonLinearSRGB: AgXFitted -> AgXLook (Punchy) -> AgXFittedEOTF
__device__ c3 film_agx(c3 c, float white_ev, float black_ev) {
    const float min_ev = dm::log2f(dm::exp2f(black_ev) * 0.18f);
    const float max_ev = dm::log2f(dm::exp2f(white_ev) * 0.18f);
    const float range = max_ev - min_ev;
    c3 r = mat3_cols(kSrgbTo2020, c);
    r = C3(fmax_(r.x, 0.0f), fmax_(r.y, 0.0f), fmax_(r.z, 0.0f));
    c3 in = mat3_cols(kInset, r);
    c3 l = C3(clampf_(dm::log2f(in.x), min_ev, max_ev), clampf_(dm::log2f(in.y), min_ev, max_ev),
              clampf_(dm::log2f(in.z), min_ev, max_ev));
    c3 v = C3(agx_curve((l.x - min_ev) / range), agx_curve((l.y - min_ev) / range), agx_curve((l.z - min_ev) / range));
    // look: luma + sat * (pow(val * slope + offset, power) - luma), slope 1, offset 0
    const float luma = dot3(v, C3(0.2126f, 0.7152f, 0.0722f));
    const float power = F_(1.05), sat = F_(1.25);
    c3 s = C3(v.x * 1.0f + 0.0f, v.y * 1.0f + 0.0f, v.z * 1.0f + 0.0f);
    c3 nv = C3(dm::powf(s.x, power), dm::powf(s.y, power), dm::powf(s.z, power));
    v = C3(luma + sat * (nv.x - luma), luma + sat * (nv.y - luma), luma + sat * (nv.z - luma));
    // EOTF: gamma 2.4 decode, outset, clamp, sRGB encode
    const float g = F_(2.4);
    c3 lin = C3(dm::powf(v.x, g), dm::powf(v.y, g), dm::powf(v.z, g));
    c3 o = mat3_cols(kOutset, lin);
    o = C3(fmax_(o.x, 0.0f), fmax_(o.y, 0.0f), fmax_(o.z, 0.0f));
    return C3(srgb_oetf(o.x), srgb_oetf(o.y), srgb_oetf(o.z));
}

// ------------------------------------------------------------ histogram
// luminanceToBin (histogram.cu:18-33)
__device__ __forceinline__ uint32_t lum_to_bin(float lum, float min_log, float inv_range) {
    if (lum < HIST_LUM_EPS) return 0u;
    const float n = clampf_((dm::log2f(lum) - min_log) * inv_range, 0.0f, 1.0f);
    return (uint32_t)((n * (float)(HIST_BINS - 2)) + 1);
}
// centerMeteringWeight (histogram.cu:35-48), radius factor 0.4
__device__ __forceinline__ float meter_weight(int W, int H, int px, int py) {
    const int cx = (int)((float)W / 2.0f), cy = (int)((float)H / 2.0f);
    const float dx = (float)(cx - px), dy = (float)(cy - py);
    float distance = __builtin_sqrtf(dx * dx + dy * dy);
    distance *= (1.0f / 0.4f);
    const float max_distance = fmin_((float)W, (float)H) / 2.0f;
    const float nd = distance / max_distance;
    const float y = clampf_((nd - 0.0f) / (1.0f - 0.0f), 0.0f, 1.0f);
    const float weight = 1.0f - (y * y * (3.0f - (2.0f * y)));
    return clampf_(weight, 0.0f, 1.0f);
}

// Temporal copy, modulation, mip-0 copy and histogram in one pass over the
// frame.  Grid-stride; one 256-lane workgroup owns 256 LDS bins.
__global__ __launch_bounds__(POST_BLOCK) void k_post_prepare(FrameBuffers fb, PostBuffers pb, PostParams p) {
    __shared__ unsigned long long s_hist[HIST_BINS];
    const int64_t npx = (int64_t)p.width * p.height;
    float white = 0.0f, black = 0.0f, inv_range = 0.0f;
    if (p.auto_exposure) {
        s_hist[threadIdx.x] = 0ull;
        white = pb.expo->white_ev;
        black = pb.expo->black_ev;
        inv_range = 1.0f / (white - black);
        __syncthreads();
    }
    for (int64_t i = (int64_t)blockIdx.x * POST_BLOCK + threadIdx.x; i < npx; i += (int64_t)gridDim.x * POST_BLOCK) {
        const float4 r = fb.render[i];
        if (p.temporal) fb.accum[i] = r;
        const uint32_t a = fb.gbuffer[i].z;
        // byteToFloat (packing.cuh:19-22); RGBSpectrum product
        const float m0 = r.x * ((float)(a & 0xFFu) / 255.0f);
        const float m1 = r.y * ((float)((a >> 8) & 0xFFu) / 255.0f);
        const float m2 = r.z * ((float)((a >> 16) & 0xFFu) / 255.0f);
        const float4 m = make_float4(m0, m1, m2, 1.0f);
        fb.render[i] = m;
        if (p.bloom && !p.upscale) pb.mip[0][i] = m;
        if (p.auto_exposure) {
            const int px = (int)(i % p.width), py = (int)(i / p.width);
            const uint32_t b = lum_to_bin(luminance(m0, m1, m2), black, inv_range);
            float w = meter_weight(p.width, p.height, px, py);
            w = (b > 0) ? w : 1.0f;
            atomicAdd(&s_hist[b], (unsigned long long)(uint32_t)(w * 8388608.0f + 0.5f));
        }
    }
    if (p.auto_exposure) {
        __syncthreads();
        const unsigned long long v = s_hist[threadIdx.x];
        if (v) atomicAdd(&pb.hist[threadIdx.x], v);
    }
}

// histogramAverageLuminanceComputeKernel + AutoExposureProgram::computeExposure
// + Renderer::setExposure, one 256-lane workgroup.  Clears the bins.
__global__ __launch_bounds__(HIST_BINS) void k_auto_exposure(PostBuffers pb, float frame_delta_ms) {
    __shared__ float s_w[HIST_BINS];
    __shared__ float s_wl[HIST_BINS];
    const uint32_t t = threadIdx.x;
    const float v = (float)((double)pb.hist[t] / 8388608.0);
    s_w[t] = v;
    s_wl[t] = v * (float)(int)t;
    pb.hist[t] = 0ull;
    __syncthreads();
    for (uint32_t cut = HIST_BINS / 2; cut > 0; cut >>= 1) {
        if (t < cut) {
            s_wl[t] += s_wl[t + cut];
            s_w[t] += s_w[t + cut];
        }
        __syncthreads();
    }
    if (t != 0) return;
    ExposureState* e = pb.expo;
    const float white = e->white_ev, black = e->black_ev;
    const float non_black = (float)__builtin_fmax((double)(s_w[0] - v), 0.001);
    const float avg_bin = s_wl[0] / non_black;
    const float range = white - black;
    const float avg_lum =
        dm::powf(2.0f, (clampf_((avg_bin - 1.0f) / (float)(HIST_BINS - 2), 0.0f, 1.0f) * range) + black);
    const float adapted = lerpf_(e->avg_lum, avg_lum, 1.0f - dm::expf(-frame_delta_ms * 0.5f));
    e->avg_lum = adapted;
    // computeExposure -> applyAperturePriority(0.1, ...)
    kpt_exposure ae = e->ae;
    float ev_target = dm::log2f((adapted * 100.0f) / 12.5f) - ae.ev_comp;
    ev_target += 6.0f;
    const float aperture = ae.aperture_f_num;
    float shutter = 1.0f / (0.1f * 1000.0f);
    const float iso = clampf_((aperture * aperture * 100.0f) / (shutter * dm::powf(2.0f, ev_target)),
                              (float)ae.min_iso, (float)ae.max_iso);
    const float ev = dm::log2f((aperture * aperture * 100.0f) / (shutter * iso));
    const float ev_diff = ev_target - ev;
    shutter = clampf_(shutter * dm::powf(2.0f, -ev_diff), ae.min_shutter_secs, ae.max_shutter_secs);
    e->ae.iso = iso;
    e->ae.shutter_speed_secs = shutter;
    // setExposure: getStandardOutputBasedExposure, film white/black
    const float l_avg = (10.0f / 0.65f) * ((aperture * aperture) / (iso * shutter));
    e->scalar = 0.18f / l_avg;
    e->white_ev = ae.white_point_ev;
    e->black_ev = ae.black_point_ev;
}

// ------------------------------------------------------------ bloom
// textureRead36Texels (bloom.cu:19-84), karisAverage (bloom.cu:6-17)
__device__ __forceinline__ c4 karis(c4 a, c4 b, c4 c, c4 d) {
    const float w0 = 1.0f / luminance(a.x, a.y, a.z);
    const float w1 = 1.0f / luminance(b.x, b.y, b.z);
    const float w2 = 1.0f / luminance(c.x, c.y, c.z);
    const float w3 = 1.0f / luminance(d.x, d.y, d.z);
    const float net = 1.0f / (w0 + w1 + w2 + w3);
    return C4((((a.x * w0 + b.x * w1) + c.x * w2) + d.x * w3) * net, (((a.y * w0 + b.y * w1) + c.y * w2) + d.y * w3) * net,
              (((a.z * w0 + b.z * w1) + c.z * w2) + d.z * w3) * net, (((a.w * w0 + b.w * w1) + c.w * w2) + d.w * w3) * net);
}
__device__ __forceinline__ c4 box4(c4 a, c4 b, c4 c, c4 d, bool k) {
    return k ? karis(a, b, c, d) : lerp4(lerp4(a, b, 0.5f), lerp4(c, d, 0.5f), 0.5f);
}

__global__ __launch_bounds__(POST_BLOCK) void k_bloom_down(Tex src, float4* dst, int dw, int dh, int use_karis) {
    const int x = (int)(blockIdx.x * 16 + (threadIdx.x & 15)), y = (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
    if (x >= dw || y >= dh) return;
    const float sx = ((float)x / (float)dw) * (float)src.w, sy = ((float)y / (float)dh) * (float)src.h;
    const bool k = use_karis != 0;
    c4 a = src.bilinear(sx + -2.0f, sy + 2.0f), b = src.bilinear(sx + 0.0f, sy + 2.0f);
    c4 c = src.bilinear(sx + 2.0f, sy + 2.0f), d = src.bilinear(sx + -2.0f, sy + 0.0f);
    c4 e = src.bilinear(sx, sy), f = src.bilinear(sx + 2.0f, sy + 0.0f);
    c4 g = src.bilinear(sx + -2.0f, sy + -2.0f), h = src.bilinear(sx + 0.0f, sy + -2.0f);
    c4 i = src.bilinear(sx + 2.0f, sy + -2.0f);
    c4 hb1 = box4(a, b, d, e, k), hb2 = box4(b, c, e, f, k), hb3 = box4(d, e, g, h, k), hb4 = box4(e, f, h, i, k);
    c4 j = src.bilinear(sx + -1.0f, sy + 1.0f), kk = src.bilinear(sx + 1.0f, sy + 1.0f);
    c4 l = src.bilinear(sx + -1.0f, sy + -1.0f), m = src.bilinear(sx + 1.0f, sy + -1.0f);
    c4 hb5 = box4(j, kk, l, m, k);
    c4 o = lerp4(hb5, lerp4(lerp4(hb1, hb2, 0.5f), lerp4(hb3, hb4, 0.5f), 0.5f), 0.5f);
    c3 oc = clamp_output3(C3(o.x, o.y, o.z));
    dst[(size_t)y * dw + x] = make_float4(oc.x, oc.y, oc.z, 1.0f);
}

__global__ __launch_bounds__(POST_BLOCK) void k_bloom_up(Tex src, float4* dst, int dw, int dh, float blend) {
    const int x = (int)(blockIdx.x * 16 + (threadIdx.x & 15)), y = (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
    if (x >= dw || y >= dh) return;
    const float sx = ((float)x / (float)dw) * (float)src.w, sy = ((float)y / (float)dh) * (float)src.h;
    // 3x3 tent (GAUSSIAN_3x3_KERNEL, bloom.cu:86-91), accumulated row by row
    const float K[3][3] = {{0.0625f, 0.125f, 0.0625f}, {0.125f, 0.25f, 0.125f}, {0.0625f, 0.125f, 0.0625f}};
    c4 acc = C4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int ty = -1; ty <= 1; ++ty) {
#pragma unroll
        for (int tx = -1; tx <= 1; ++tx) {
            const c4 c = src.bilinear(sx + (float)tx, sy + (float)ty);
            const float k = K[ty + 1][tx + 1];
            acc = C4(acc.x + k * c.x, acc.y + k * c.y, acc.z + k * c.z, acc.w + k * c.w);
        }
    }
    const size_t di = (size_t)y * dw + x;
    acc = lerp4(ld(dst[di]), acc, blend);
    c3 oc = clamp_output3(C3(acc.x, acc.y, acc.z));
    dst[di] = make_float4(oc.x, oc.y, oc.z, 1.0f);
}

// ------------------------------------------------------------ upscale
// upscale (upscale.cuh:260-298, the build without USE_FSR): output pixel
// (x, y) reads the render buffer bilinearly (textureReadBilinear with alpha
// filtered) at (x / out_width * width, y / out_height * height); the result
// is the output texture that bloom (as its mip 0) and the post pass read.
__global__ __launch_bounds__(POST_BLOCK) void k_upscale(FrameBuffers fb, PostBuffers pb, PostParams p) {
    const int64_t npx = (int64_t)p.out_width * p.out_height;
    const int64_t i = (int64_t)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (i >= npx) return;
    const int x = (int)(i % p.out_width), y = (int)(i / p.out_width);
    const float u = (float)x / (float)p.out_width, v = (float)y / (float)p.out_height;
    const Tex src{fb.render, p.width, p.height};
    const float4 c = st(src.bilinear(u * (float)p.width, v * (float)p.height));
    pb.up[i] = c;
    if (p.bloom) pb.mip[0][i] = c;
}

// ------------------------------------------------------------ FSR 1.0
// The reference's USE_FSR build of the upscale (upscale.cuh:13-258 EASU,
// 302-378 RCAS; kernels.cu:84-103 runs EASU render -> output and RCAS output
// -> back buffer, renderer.cu:138-147 copies the back buffer to the output).
// Here EASU writes pb.back (free until the post pass) and RCAS writes the
// upscaled output texture pb.up (and bloom's mip 0).  Float vector
// arithmetic as vector_types_extension.cuh defines it, the reference's double
// literals promoted as written, inversesqrt taken as the IEEE 1/sqrtf, RCAS
// with sharpness 0 (con = 1) and no denoise term; one thread per output
// pixel, each tap a bilinear read of the clamped source.
struct f2 {
    float x, y;
};
__device__ __forceinline__ c3 easu_cf(float px, float py, const Tex& t) {   // textureReadBilinearUV(p, true)
    const c4 c = t.bilinear(px * (float)t.w, py * (float)t.h);
    return C3(c.x, c.y, c.z);
}
__device__ __forceinline__ float easu_luma(c3 c) { return (float)((double)c.y + 0.5 * (double)(c.x + c.z)); }
__device__ __forceinline__ float hm_clamp(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }
// FsrEasuSetF (upscale.cuh:103-130)
__device__ __forceinline__ void easu_set(f2& dir, float& len, float w, float lA, float lB, float lC, float lD,
                                         float lE) {
    float lenX = fmaxf(fabsf(lD - lC), fabsf(lC - lB));
    const float dirX = lD - lB;
    dir.x += dirX * w;
    lenX = hm_clamp(fabsf(dirX) / lenX, 0.0f, 1.0f);
    lenX *= lenX;
    len += lenX * w;
    float lenY = fmaxf(fabsf(lE - lC), fabsf(lC - lA));
    const float dirY = lE - lA;
    dir.y += dirY * w;
    lenY = hm_clamp(fabsf(dirY) / lenY, 0.0f, 1.0f);
    lenY *= lenY;
    len += lenY * w;
}
// FsrEasuTapF (upscale.cuh:64-101)
__device__ __forceinline__ void easu_tap(c3& aC, float& aW, f2 off, f2 dir, f2 len, float lob, float clp, c3 c) {
    float vx = off.x * dir.x + off.y * dir.y, vy = off.x * -dir.y + off.y * dir.x;
    vx *= len.x;
    vy *= len.y;
    const float d2 = fminf(vx * vx + vy * vy, clp);
    float wB = (float)(0.4 * (double)d2 - 1.0);
    float wA = (float)((double)(lob * d2) - 1.0);
    wB *= wB;
    wA *= wA;
    wB = (float)(1.5625 * (double)wB - 0.5625);
    const float w = wB * wA;
    aC = C3(aC.x + c.x * w, aC.y + c.y * w, aC.z + c.z * w);
    aW += w;
}

// FsrEasuCon + FsrEasuF (upscale.cuh:18-62, 132-258): render (width x height)
// -> pb.back (out_width x out_height).
__global__ __launch_bounds__(POST_BLOCK) void k_fsr_easu(FrameBuffers fb, PostBuffers pb, PostParams p) {
    const int64_t npx = (int64_t)p.out_width * p.out_height;
    const int64_t i = (int64_t)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (i >= npx) return;
    const int x = (int)(i % p.out_width), y = (int)(i / p.out_width);
    const Tex t{fb.render, p.width, p.height};
    const float iw = (float)p.width, ih = (float)p.height, ow = (float)p.out_width, oh = (float)p.out_height;
    const float c0x = iw / ow, c0y = ih / oh;
    const float c0z = (float)(0.5 * (double)iw / (double)ow - 0.5), c0w = (float)(0.5 * (double)ih / (double)oh - 0.5);
    const float c1x = 1.0f / iw, c1y = 1.0f / ih, c1z = 1.0f / iw, c1w = -1.0f / ih;
    const float c2x = -1.0f / iw, c2y = 2.0f / ih, c2z = 1.0f / iw, c2w = 2.0f / ih;
    const float c3x = 0.0f / iw, c3y = 4.0f / ih;
    f2 pp{(float)x * c0x + c0z, (float)y * c0y + c0w};
    const f2 fp{floorf(pp.x), floorf(pp.y)};
    pp = f2{pp.x - fp.x, pp.y - fp.y};
    const f2 p0{fp.x * c1x + c1z, fp.y * c1y + c1w};
    const f2 p1{p0.x + c2x, p0.y + c2y}, p2{p0.x + c2z, p0.y + c2w}, p3{p0.x + c3x, p0.y + c3y};
    const float ox = -0.5f * c1x, oy = 0.5f * c1x, oz = -0.5f * c1y, ow_ = 0.5f * c1y;
    const c3 bC = easu_cf(p0.x + ox, p0.y + ow_, t), cC = easu_cf(p0.x + oy, p0.y + ow_, t);
    const c3 iC = easu_cf(p1.x + ox, p1.y + ow_, t), jC = easu_cf(p1.x + oy, p1.y + ow_, t);
    const c3 fC = easu_cf(p1.x + oy, p1.y + oz, t), eC = easu_cf(p1.x + ox, p1.y + oz, t);
    const c3 kC = easu_cf(p2.x + ox, p2.y + ow_, t), lC = easu_cf(p2.x + oy, p2.y + ow_, t);
    const c3 hC = easu_cf(p2.x + oy, p2.y + oz, t), gC = easu_cf(p2.x + ox, p2.y + oz, t);
    const c3 oC = easu_cf(p3.x + oy, p3.y + oz, t), nC = easu_cf(p3.x + ox, p3.y + oz, t);
    const float bL = easu_luma(bC), cL = easu_luma(cC), iL = easu_luma(iC), jL = easu_luma(jC),
                fL = easu_luma(fC), eL = easu_luma(eC), kL = easu_luma(kC), lL = easu_luma(lC),
                hL = easu_luma(hC), gL = easu_luma(gC), oL = easu_luma(oC), nL = easu_luma(nC);
    f2 dir{0.0f, 0.0f};
    float len = 0.0f;
    easu_set(dir, len, (float)((1.0 - (double)pp.x) * (1.0 - (double)pp.y)), bL, eL, fL, gL, jL);
    easu_set(dir, len, (float)((double)pp.x * (1.0 - (double)pp.y)), cL, fL, gL, hL, kL);
    easu_set(dir, len, (float)((1.0 - (double)pp.x) * (double)pp.y), fL, iL, jL, kL, nL);
    easu_set(dir, len, pp.x * pp.y, gL, jL, kL, lL, oL);
    float dirR = dir.x * dir.x + dir.y * dir.y;
    const bool zro = (double)dirR < (1.0 / 32768.0);
    dirR = 1.0f / sqrtf(dirR);
    dirR = zro ? 1.0f : dirR;
    dir.x = zro ? 1.0f : dir.x;
    dir = f2{dir.x * dirR, dir.y * dirR};
    len = (float)((double)len * 0.5);
    len *= len;
    const float stretch = (dir.x * dir.x + dir.y * dir.y) / fmaxf(fabsf(dir.x), fabsf(dir.y));
    const f2 len2{(float)(1.0 + ((double)stretch - 1.0) * (double)len), (float)(1.0 - 0.5 * (double)len)};
    const float lob = (float)(0.5 - 0.29 * (double)len);
    const float clp = (float)(1.0 / (double)lob);
    const c3 min4 = C3(fminf(fminf(fC.x, gC.x), fminf(jC.x, kC.x)), fminf(fminf(fC.y, gC.y), fminf(jC.y, kC.y)),
                       fminf(fminf(fC.z, gC.z), fminf(jC.z, kC.z)));
    const c3 max4 = C3(fmaxf(fmaxf(fC.x, gC.x), fmaxf(jC.x, kC.x)), fmaxf(fmaxf(fC.y, gC.y), fmaxf(jC.y, kC.y)),
                       fmaxf(fmaxf(fC.z, gC.z), fmaxf(jC.z, kC.z)));
    c3 aC = C3(0.0f, 0.0f, 0.0f);
    float aW = 0.0f;
    easu_tap(aC, aW, f2{0.0f - pp.x, -1.0f - pp.y}, dir, len2, lob, clp, bC);
    easu_tap(aC, aW, f2{1.0f - pp.x, -1.0f - pp.y}, dir, len2, lob, clp, cC);
    easu_tap(aC, aW, f2{-1.0f - pp.x, 1.0f - pp.y}, dir, len2, lob, clp, iC);
    easu_tap(aC, aW, f2{0.0f - pp.x, 1.0f - pp.y}, dir, len2, lob, clp, jC);
    easu_tap(aC, aW, f2{0.0f - pp.x, 0.0f - pp.y}, dir, len2, lob, clp, fC);
    easu_tap(aC, aW, f2{-1.0f - pp.x, 0.0f - pp.y}, dir, len2, lob, clp, eC);
    easu_tap(aC, aW, f2{1.0f - pp.x, 1.0f - pp.y}, dir, len2, lob, clp, kC);
    easu_tap(aC, aW, f2{2.0f - pp.x, 1.0f - pp.y}, dir, len2, lob, clp, lC);
    easu_tap(aC, aW, f2{2.0f - pp.x, 0.0f - pp.y}, dir, len2, lob, clp, hC);
    easu_tap(aC, aW, f2{1.0f - pp.x, 0.0f - pp.y}, dir, len2, lob, clp, gC);
    easu_tap(aC, aW, f2{1.0f - pp.x, 2.0f - pp.y}, dir, len2, lob, clp, oC);
    easu_tap(aC, aW, f2{0.0f - pp.x, 2.0f - pp.y}, dir, len2, lob, clp, nC);
    const float rx = aC.x / aW, ry = aC.y / aW, rz = aC.z / aW;
    pb.back[i] = make_float4(fminf(max4.x, fmaxf(min4.x, rx)), fminf(max4.y, fmaxf(min4.y, ry)),
                             fminf(max4.z, fmaxf(min4.z, rz)), 1.0f);
}

// FsrRcasF with con = exp2(-0) = 1 (upscale.cuh:316-378, 389-408): pb.back ->
// pb.up (and bloom's mip 0), output size.
__global__ __launch_bounds__(POST_BLOCK) void k_fsr_rcas(PostBuffers pb, PostParams p) {
    const int64_t npx = (int64_t)p.out_width * p.out_height;
    const int64_t i = (int64_t)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (i >= npx) return;
    const int x = (int)(i % p.out_width), y = (int)(i / p.out_width);
    const Tex t{pb.back, p.out_width, p.out_height};
    const float w = (float)p.out_width, h = (float)p.out_height;
    auto load = [&](float px, float py) {   // textureReadBilinearUV(p / dimensions, true)
        const float u = px / w, v = py / h;
        const c4 c = t.bilinear(u * w, v * h);
        return C3(c.x, c.y, c.z);
    };
    const float sx = (float)x, sy = (float)y;
    const c3 b = load(sx + 0.0f, sy + -1.0f), d = load(sx + -1.0f, sy + 0.0f), e = load(sx, sy),
             f = load(sx + 1.0f, sy + 0.0f), hh = load(sx + 0.0f, sy + 1.0f);
    const c3 mn4 = C3(fminf(b.x, fminf(f.x, hh.x)), fminf(b.y, fminf(f.y, hh.y)), fminf(b.z, fminf(f.z, hh.z)));
    const c3 mx4 = C3(fmaxf(b.x, fmaxf(f.x, hh.x)), fmaxf(b.y, fmaxf(f.y, hh.y)), fmaxf(b.z, fmaxf(f.z, hh.z)));
    const c3 hitMin = C3(mn4.x / (4.0f * mx4.x), mn4.y / (4.0f * mx4.y), mn4.z / (4.0f * mx4.z));
    const c3 hitMax = C3((1.0f - mx4.x) / (4.0f * mn4.x + -4.0f), (1.0f - mx4.y) / (4.0f * mn4.y + -4.0f),
                         (1.0f - mx4.z) / (4.0f * mn4.z + -4.0f));
    const float m = fmaxf(fmaxf(-hitMin.x, hitMax.x), fmaxf(fmaxf(-hitMin.y, hitMax.y), fmaxf(-hitMin.z, hitMax.z)));
    const float lobe = (float)(fmax(-(0.25 - (1.0 / 16.0)), fmin((double)m, 0.0)) * 1.0);
    const c3 s = C3(((b.x + d.x) + hh.x) + f.x, ((b.y + d.y) + hh.y) + f.y, ((b.z + d.z) + hh.z) + f.z);
    const float den = (float)(4.0 * (double)lobe + 1.0);
    const float4 o = make_float4((lobe * s.x + e.x) / den, (lobe * s.y + e.y) / den, (lobe * s.z + e.z) / den, 1.0f);
    pb.up[i] = o;
    if (p.bloom) pb.mip[0][i] = o;
}

// ------------------------------------------------------------ post
// computePostProcess (kernels.cu:475-504): bloom blend, exposure in Yxy, AgX,
// at the output size over the output texture (the render buffer, or its
// upscale).
// Into the back buffer when the effects pass follows, else into the output.
__global__ __launch_bounds__(POST_BLOCK) void k_post_final(FrameBuffers fb, PostBuffers pb, PostParams p) {
    const int64_t npx = (int64_t)p.out_width * p.out_height;
    const int64_t i = (int64_t)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (i >= npx) return;
    const ExposureState* e = pb.expo;
    const float4 r = p.upscale ? pb.up[i] : fb.render[i];
    c3 c = C3(r.x, r.y, r.z);
    if (p.bloom) {
        const float4 b = pb.mip[0][i];
        const float t = p.bloom_final_blend;
        c = C3(lerpf_(c.x, b.x, t), lerpf_(c.y, b.y, t), lerpf_(c.z, b.z, t));
    }
    c3 yxy = xyz_to_Yxy(rgb_to_xyz(c));
    yxy.x *= e->scalar;
    c = xyz_to_rgb(Yxy_to_xyz(yxy));
    const c3 o = film_agx(c, e->white_ev, e->black_ev);
    (p.effects ? pb.back : fb.output)[i] = make_float4(o.x, o.y, o.z, 1.0f);
}

// ------------------------------------------------------------ effects
// Float to int as the device's conversion does it (round toward zero,
// saturating, NaN to 0), the make_int2(float2) of the texture accessors.
__device__ __forceinline__ int f2i_sat(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int)x;
}

// getFishEyeUV (kernels.cu:340-376) with power 1.65: the fisheye branch.
// 0.5 / prop is a double division narrowed to float (make_float2 of a double
// expression); normalize is v * (1 / sqrt(dot(v, v))), so the centre pixel,
// where d = 0, gets NaN coordinates as the reference's does.
__device__ __forceinline__ void fisheye_uv(float px, float py, float fw, float fh, float* u, float* v) {
    const float p_x = px / fw, p_y = py / fw;
    const float prop = fw / fh;
    const float m_x = 0.5f, m_y = (float)(0.5 / (double)prop);
    const float d_x = p_x - m_x, d_y = p_y - m_y;
    const float r = __builtin_sqrtf((d_x * d_x) + (d_y * d_y));
    const float power = 1.65f;
    const float bind = __builtin_sqrtf((m_x * m_x) + (m_y * m_y));
    const float inv = 1.0f / __builtin_sqrtf((d_x * d_x) + (d_y * d_y));
    const float n_x = d_x * inv, n_y = d_y * inv;
    const float t = dm::tanf(r * power);
    const float tb = dm::tanf(bind * power);
    const float a_x = n_x * t, a_y = n_y * t;
    const float b_x = a_x * bind, b_y = a_y * bind;
    const float c_x = b_x / tb, c_y = b_y / tb;
    *u = m_x + c_x;
    *v = (m_y + c_y) * prop;
}

// computeEffects (kernels.cu:508-528): fisheye coordinates, then R, G and B
// point-sampled (textureReadNearestUV: truncate, clamp to the edge) at the
// coordinates plus, at, and minus a radial offset.
__global__ __launch_bounds__(POST_BLOCK) void k_fx(FrameBuffers fb, PostBuffers pb, PostParams p) {
    const int64_t npx = (int64_t)p.out_width * p.out_height;
    const int64_t i = (int64_t)blockIdx.x * POST_BLOCK + threadIdx.x;
    if (i >= npx) return;
    const int W = p.out_width, H = p.out_height;
    const int x = (int)(i % W), y = (int)(i / W);
    const float fw = (float)W, fh = (float)H;
    const float uc = ((float)x + 0.5f) / fw, vc = ((float)y + 0.5f) / fh;   // getShadingJob uv
    float u, v;
    fisheye_uv((float)x, (float)y, fw, fh, &u, &v);
    const float o_x = dm::powf(__builtin_fabsf(uc - 0.5f) * 2.0f, 2.0f) * 0.01f;
    const float o_y = dm::powf(__builtin_fabsf(vc - 0.5f) * 2.0f, 2.0f) * 0.01f;
    auto tap = [&](float tu, float tv) {
        int ix = f2i_sat(tu * fw), iy = f2i_sat(tv * fh);
        ix = ix < 0 ? 0 : (ix > W - 1 ? W - 1 : ix);
        iy = iy < 0 ? 0 : (iy > H - 1 ? H - 1 : iy);
        return pb.back[(size_t)iy * W + ix];
    };
    const float r = tap(u + o_x, v + o_y).x;
    const float g = tap(u, v).y;
    const float b = tap(u - o_x, v - o_y).z;
    // textureWriteUV at the pixel's own uv (zero boundary: out of range dropped)
    const int wx = f2i_sat(uc * fw), wy = f2i_sat(vc * fh);
    if (wx >= 0 && wx < W && wy >= 0 && wy < H) fb.output[(size_t)wy * W + wx] = make_float4(r, g, b, 1.0f);
}

}  // namespace

int bloom_levels(int width, int height) {
    auto flog2 = [](int v) { int l = -1; while (v) { v >>= 1; ++l; } return l; };
    int l = flog2(width) < flog2(height) ? flog2(width) : flog2(height);
    return l < MAX_MIPS - 1 ? l : MAX_MIPS - 1;
}

hipError_t launch_post_frame(hipStream_t s, const FrameBuffers& fb, const PostBuffers& pb, const PostParams& p,
                             int num_cus) {
    const int64_t npx = (int64_t)p.width * p.height;
    if (npx <= 0) return hipSuccess;
    const int64_t blocks = (npx + POST_BLOCK - 1) / POST_BLOCK;
    const unsigned pgrid = (unsigned)std::min<int64_t>(blocks, (int64_t)std::max(1, num_cus) * 8);
    hipLaunchKernelGGL(k_post_prepare, dim3(pgrid), dim3(POST_BLOCK), 0, s, fb, pb, p);
    if (p.auto_exposure) hipLaunchKernelGGL(k_auto_exposure, dim3(1), dim3(HIST_BINS), 0, s, pb, p.frame_delta_ms);
    // from here on the output size: bloom, post and effects
    const int64_t onpx = (int64_t)p.out_width * p.out_height;
    const unsigned oblocks = (unsigned)((onpx + POST_BLOCK - 1) / POST_BLOCK);
    if (p.upscale && p.fsr) {
        hipLaunchKernelGGL(k_fsr_easu, dim3(oblocks), dim3(POST_BLOCK), 0, s, fb, pb, p);
        hipLaunchKernelGGL(k_fsr_rcas, dim3(oblocks), dim3(POST_BLOCK), 0, s, pb, p);
    } else if (p.upscale) {
        hipLaunchKernelGGL(k_upscale, dim3(oblocks), dim3(POST_BLOCK), 0, s, fb, pb, p);
    }
    const int OW = p.out_width, OH = p.out_height;
    if (p.bloom) {
        for (int k = 0; k < pb.levels; ++k) {
            const int sw = OW >> k, sh = OH >> k, dw = OW >> (k + 1), dh = OH >> (k + 1);
            Tex src{pb.mip[k], sw, sh};
            hipLaunchKernelGGL(k_bloom_down, dim3((dw + 15) / 16, (dh + 15) / 16), dim3(POST_BLOCK), 0, s, src,
                               pb.mip[k + 1], dw, dh, (p.karis && k == 0) ? 1 : 0);
        }
        for (int k = pb.levels; k > 0; --k) {
            const int sw = OW >> k, sh = OH >> k, dw = OW >> (k - 1), dh = OH >> (k - 1);
            Tex src{pb.mip[k], sw, sh};
            hipLaunchKernelGGL(k_bloom_up, dim3((dw + 15) / 16, (dh + 15) / 16), dim3(POST_BLOCK), 0, s, src,
                               pb.mip[k - 1], dw, dh, p.bloom_internal_blend);
        }
    }
    hipLaunchKernelGGL(k_post_final, dim3(oblocks), dim3(POST_BLOCK), 0, s, fb, pb, p);
    if (p.effects) hipLaunchKernelGGL(k_fx, dim3(oblocks), dim3(POST_BLOCK), 0, s, fb, pb, p);
    return hipGetLastError();
}

}  // namespace kpt
