// kpt_shade.h — surface, BSDF, sky and sampling routines of the path tracer
// for gfx950 (device code).  Each routine names the reference routine it
// computes; expression association follows the reference so results are
// bit-identical to the IEEE restatement (kpt_dmath.h transcendentals, no FMA
// contraction).
#pragma once
#include "kpt_trace.h"

namespace kpt {

#define KPT_PI 3.14159265358979323846f
#define KPT_INV_PI 0.31830988618379067154f
#define KPT_INV_U32 (1.0f / 4294967296.0f)
#define KPT_INV_255 0.00392156862f
#define KPT_HIT_EPS 0.001f
#define KPT_SOLAR_DIST 149.597e9f
#define KPT_SUN_NITS 6.0e5f
#define KPT_EARTH_R 6378137.0f
#define KPT_ATMOS_R 6438137.0f

// ---------------------------------------------------------------- sampler
// RNG::pcg_hash, samplers.cuh:12-17
KD uint32_t pcg_hash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
KD float rng1(uint32_t& seed) { seed = pcg_hash(seed); return (float)seed * KPT_INV_U32; }
// IndependentSampler::initPixelSeed, samplers.cuh:24-29 (mode 0), or the
// pixel-index-xor-frame policy (mode 1).
KD uint32_t init_seed(int px, int py, int width, int frame, int mode) {
    if (mode == 1) return (uint32_t)(px + py * width) ^ (uint32_t)frame;
    uint32_t s = (uint32_t)(px + py * width);
    s = (uint32_t)((float)s * ((frame == 0) ? 1.0f : (float)frame));
    return s;
}

// ---------------------------------------------------------------- BSDF
// GGX, bsdf.cuh:18-88
KD float ggx_D(float alpha, v3 wm) {
    float a = alpha;
    if (alpha < 0.001f) a = fmaxf_(a, 1.0e-3f);
    float a2 = sqr(a);
    float ndh = sqr(saturate(wm.z));
    float b = (float)((double)ndh * ((double)a2 - 1.0) + 1.0);
    return a2 / (KPT_PI * sqr(b));
}
KD float ggx_G(float alpha, v3 wo, v3 wi) {
    float nv = fmaxf_(wo.z, 0.0f);
    float nl = __builtin_fabsf(wi.z);
    float tv = __builtin_sqrtf(sqr(alpha) + (1.0f - sqr(alpha)) * sqr(nv));
    float tl = __builtin_sqrtf(sqr(alpha) + (1.0f - sqr(alpha)) * sqr(nl));
    float g1v = (2.0f * nv) / (nv + tv);
    float g1l = (2.0f * nl) / (nl + tl);
    return g1v * g1l;
}
KD v3 ggx_sample_wm(float alpha, float e0, float e1) {
    float a2 = sqr(alpha);
    float theta = dm::acosf(__builtin_sqrtf((1.0f - e0) / ((a2 - 1.0f) * e0 + 1.0f)));
    float phi = 2.0f * KPT_PI * e1;
    v3 wm = mk3(dm::sinf(theta) * dm::cosf(phi), dm::sinf(theta) * dm::sinf(phi), dm::cosf(theta));
    return normalize(wm);
}
// fresnelDielectric, bsdf.cuh:122-141
KD float fresnel_dielectric(float c, float ior) {
    c = clampf(c, -1.0f, 1.0f);
    if (c < 0.0f) { ior = 1.0f / ior; c = -c; }
    float s2 = (1.0f - c * c);
    float s2t = s2 / (ior * ior);
    if (s2t >= 1.0f) return 1.0f;
    float ct = __builtin_sqrtf(1.0f - s2t);
    float rp = (ior * c - ct) / (ior * c + ct);
    float rs = (c - ior * ct) / (c + ior * ct);
    return (rp * rp + rs * rs) * 0.5f;
}
// fresnelSchlick, bsdf.cuh:144-148
KD v3 fresnel_schlick(float c, v3 F0) {
    v3 F = F0 + (1.0f - F0) * pow5(1.0f - c);
    return clamp3(F, mk3(0.0f, 0.0f, 0.0f), mk3(1.0f, 1.0f, 1.0f));
}
// microfacetBRDF, bsdf.cuh:151-164
KD float microfacet_brdf(float alpha, v3 wo, v3 wi, v3 wm) {
    float nv = saturate(wo.z);
    float nl = saturate(wi.z);
    float D = ggx_D(alpha, wm);
    float G = ggx_G(alpha, wo, wi);
    float f = (D * G) / (4.0f * fmaxf_(nv, 1.0e-3f) * nl);
    if (wi.z <= 0.0f || dot(wi, wm) <= 0.0f) f = 0.0f;
    return f;
}
// refract, vector_maths.cuh:46-65
KD bool refract_(v3 wi, v3 n, float ior, v3* wt) {
    float c = dot(wi, n);
    if (c < 0.0f) { ior = 1.0f / ior; c *= -1.0f; n = n * -1.0f; }
    float s2 = (1.0f - c * c);
    float s2t = s2 / (ior * ior);
    if (s2t >= 1.0f) return false;
    float ct = __builtin_sqrtf(1.0f - s2t);
    *wt = (-1.0f * wi) / ior + (c / ior - ct) * n;
    return true;
}
KD bool same_hemi_z(v3 a, v3 b) { return (a.z * b.z) > 0.0f; }   // sameHemisphere(a, b, (0,0,1))

enum { SC_ABSORBED = 0, SC_REFLECTED = 2, SC_TRANSMITTED = 4, SC_DIFFUSE = 8, SC_GLOSSY = 16, SC_SPECULAR = 32 };

// UnifiedBSDF, bsdf.cuh:166-624.  basis columns: tangent, bitangent, normal.
struct BSDF {
    m3 basis;
    v3 albedo;
    float alpha;        // GGX alpha = roughness^2 (possibly regularised)
    float roughness;
    float metallic;
    float transmission;
    float ior;
    bool backface;
};

struct BSample {
    int scatter;
    v3 f;
    v3 wi;
    float pdf;
};

KD float f_diffuse(v3 wo, v3 wi) { float s = (wo.z * wi.z > 0.0f) ? 1.0f : 0.0f; return s * KPT_INV_PI; }
KD float pdf_diffuse(v3 wi) { return wi.z * KPT_INV_PI; }
KD float pdf_glossy(const BSDF& b, v3 wo, v3 wm) {
    float voh = fmaxf_(dot(wo, wm), 0.0f);
    float noh = fmaxf_(wm.z, 0.0f);
    float D = ggx_D(b.alpha, wm);
    return (D * noh) / (4.0f * voh);
}
KD float f_glossy(const BSDF& b, v3 wo, v3 wi, v3 wm) {
    float M = microfacet_brdf(b.alpha, wo, wi, wm);
    float F = fresnel_dielectric(dot(wo, wm), b.ior);
    return F * M;
}
KD float pdf_btdf(const BSDF& b, v3 wo, v3 wi, v3 ht, float eta) {
    float temp = dot(wi, ht) * eta + dot(wo, ht);
    float dw = absdot(wi, ht) / (temp * temp);
    float noh = __builtin_fabsf(ht.z);
    float D = ggx_D(b.alpha, ht);
    float F = fresnel_dielectric(absdot(wo, ht), eta);
    return (D * noh) * dw * (1.0f - F);
}
KD v3 f_btdf(const BSDF& b, v3 wo, v3 wi, v3 ht, float eta) {
    float temp = dot(wi, ht) * eta + dot(wo, ht);
    float F = fresnel_dielectric(absdot(wo, ht), eta);
    float T = ggx_D(b.alpha, ht) * ggx_G(b.alpha, wo, wi) *
              __builtin_fabsf(dot(wi, ht) * dot(wo, ht) / (wi.z * wo.z * temp * temp));
    return b.albedo * (1.0f - F) * T;
}
KD v3 f_opaque(const BSDF& b, v3 wo, v3 wi) {
    v3 wm = normalize(wo + wi);
    float g = f_glossy(b, wo, wi, wm);
    v3 gv = mk3(g, g, g);
    v3 wd = 1.0f - gv;
    v3 diff = wd * (b.albedo * f_diffuse(wo, wi));
    return diff + gv;
}
KD float pdf_opaque(const BSDF& b, v3 wo, v3 wi) {
    v3 wm = normalize(wo + wi);
    float ps = pdf_glossy(b, wo, wm);
    float pd = pdf_diffuse(wi);
    return 0.5f * (ps + pd);
}
KD v3 f_conductor(const BSDF& b, v3 wo, v3 wi) {
    v3 wm = normalize(wo + wi);
    float nov = clampf(wo.z, 0.0f, 1.0f);
    float nol = clampf(wi.z, 0.0f, 1.0f);
    if (nov == 0.0f || nol == 0.0f) return mk3(0.0f, 0.0f, 0.0f);
    float voh = clampf(dot(wo, wm), 0.0f, 1.0f);
    v3 M = mk3(0.0f, 0.0f, 0.0f);
    if (nol > 0.0f && voh > 0.0f) {
        v3 F = fresnel_schlick(voh, b.albedo);
        M = F * microfacet_brdf(b.alpha, wo, wi, wm);
    }
    return M;
}
KD float pdf_conductor(const BSDF& b, v3 wo, v3 wi) { v3 wm = normalize(wo + wi); return pdf_glossy(b, wo, wm); }
KD v3 f_transparent(const BSDF& b, v3 wo, v3 wi) {
    float co = wo.z, ci = wi.z;
    bool refl = (co * ci) > 0.0f;
    float eta = 1.0f;
    if (!refl) eta = (!b.backface) ? b.ior : (1.0f / b.ior);
    v3 wm = (eta * wi + wo);
    wm = normalize((wm.z > 0.0f) ? wm : -wm);
    if ((dot(wm, wi) * ci < 0.0f) || (dot(wm, wo) * co < 0.0f)) return mk3(0.0f, 0.0f, 0.0f);
    float F = fresnel_dielectric(__builtin_fabsf(dot(wo, wm)), b.ior);
    float T = 1.0f - F;
    if (refl) { float v = F * microfacet_brdf(b.alpha, wo, wi, wm); return mk3(v, v, v); }
    float temp = dot(wi, wm) * eta + dot(wo, wm);
    float dw = __builtin_fabsf(dot(wi, wm)) * __builtin_fabsf(dot(wo, wm)) / (temp * temp);
    float Tss = ggx_D(b.alpha, wm) * ggx_G(b.alpha, wo, wi) * dw / (__builtin_fabsf(ci * co));
    return T * b.albedo * Tss;
}
KD float pdf_transparent(const BSDF& b, v3 wo, v3 wi) {
    float co = wo.z, ci = wi.z;
    bool refl = co * ci > 0.0f;
    float eta = 1.0f;
    if (!refl) eta = (!b.backface) ? b.ior : 1.0f / b.ior;
    v3 wm = (eta * wi + wo);
    wm = normalize((wm.z > 0.0f) ? wm : -wm);
    if (dot(wm, wi) * ci < 0.0f || dot(wm, wo) * co < 0.0f) return 0.0f;
    float F = fresnel_dielectric(__builtin_fabsf(dot(wo, wm)), b.ior);
    float T = 1.0f - F;
    if (refl) return pdf_glossy(b, wo, wm) * F;
    float temp = dot(wi, wm) + dot(wo, wm) / eta;
    float dw = __builtin_fabsf(dot(wo, wm)) / (temp * temp);
    float noh = fmaxf_(wm.z, 0.0f);
    float D = ggx_D(b.alpha, wm);
    return D * noh * dw * T;
}
KD v3 bsdf_f(const BSDF& b, v3 wwo, v3 wwi) {
    v3 wo = mul3t(b.basis, wwo);
    v3 wi = mul3t(b.basis, wwi);
    float wm = b.metallic;
    float wt = (1.0f - wm) * b.transmission;
    float wd = (1.0f - wm) * (1.0f - b.transmission);
    v3 F = mk3(0.0f, 0.0f, 0.0f);
    if (wm > 0.0f) F = F + (f_conductor(b, wo, wi) * wm);
    if (wt > 0.0f) F = F + (f_transparent(b, wo, wi) * wt);
    if (wd > 0.0f) F = F + (f_opaque(b, wo, wi) * wd);
    return F;
}
KD float bsdf_pdf(const BSDF& b, v3 wwo, v3 wwi) {
    v3 wi = mul3t(b.basis, wwi);
    v3 wo = mul3t(b.basis, wwo);
    float wm = b.metallic;
    float wt = (1.0f - wm) * b.transmission;
    float wd = (1.0f - wm) * (1.0f - b.transmission);
    float p = 0.0f;
    if (wm > 0.0f) p += wm * pdf_conductor(b, wo, wi);
    if (wt > 0.0f) p += wt * pdf_transparent(b, wo, wi);
    if (wd > 0.0f) p += wd * pdf_opaque(b, wo, wi);
    return p;
}
KD BSample bsdf_sample(const BSDF& b, v3 wwo, v2 u, v2 X2) {
    v3 wo = mul3t(b.basis, wwo);
    float pm = b.metallic;
    float pt = (1.0f - pm) * b.transmission;
    BSample bs;
    if (X2.x < pm) {
        // sampleConductor, bsdf.cuh:355-365
        v3 wm = ggx_sample_wm(b.alpha, u.x, u.y);
        v3 wi = reflect(-wo, wm);
        bs.scatter = SC_GLOSSY | SC_REFLECTED;
        bs.f = f_conductor(b, wo, wi);
        bs.wi = wi;
        bs.pdf = pdf_glossy(b, wo, wm);
    } else if (X2.x < pt) {
        // sampleTransparentDielectric, bsdf.cuh:394-450 (non-cumulative test, :256)
        const bool inside = b.backface;
        float eta = inside ? (1.0f / b.ior) : b.ior;
        v3 ht = ggx_sample_wm(b.alpha, u.x, u.y);
        v3 wi = mk3(0.0f, 0.0f, 0.0f);
        bool tir = !refract_(wo, ht, eta, &wi);
        const float gp = 0.5f;
        if ((X2.y < gp && !inside) || (tir && inside)) {
            wi = reflect(-wo, ht);
            if (!same_hemi_z(wi, wo)) {
                bs.scatter = SC_ABSORBED; bs.f = mk3(0.0f, 0.0f, 0.0f); bs.wi = mk3(0.0f, 0.0f, 0.0f); bs.pdf = 0.0f;
            } else {
                float M = microfacet_brdf(b.alpha, wo, wi, ht);
                float F = fresnel_dielectric(dot(wo, ht), eta);
                float v = F * M;
                v3 f = mk3(v, v, v);
                if (!tir) f = f * (1.0f / gp);
                bs.scatter = SC_REFLECTED | SC_GLOSSY; bs.f = f; bs.wi = wi; bs.pdf = pdf_glossy(b, wo, ht);
            }
        } else if (tir || same_hemi_z(wi, wo) || wi.z == 0.0f) {
            bs.scatter = SC_ABSORBED; bs.f = mk3(0.0f, 0.0f, 0.0f); bs.wi = mk3(0.0f, 0.0f, 0.0f); bs.pdf = 0.0f;
        } else {
            float pdf = pdf_btdf(b, wo, wi, ht, eta);
            v3 f = f_btdf(b, wo, wi, ht, eta);
            if (!inside) f = f * (1.0f / (1.0f - gp));
            bs.scatter = SC_TRANSMITTED | SC_GLOSSY; bs.f = f; bs.wi = wi; bs.pdf = pdf;
        }
    } else {
        // sampleOpaqueDielectric, bsdf.cuh:298-329
        const float gp = 0.5f;
        if (X2.y < gp) {
            v3 wm = ggx_sample_wm(b.alpha, u.x, u.y);
            v3 wi = reflect(-wo, wm);
            float g = f_glossy(b, wo, wi, wm);
            v3 f = mk3(g, g, g);
            f = f * (1.0f / gp);
            bs.scatter = SC_GLOSSY | SC_REFLECTED; bs.f = f; bs.wi = wi; bs.pdf = pdf_glossy(b, wo, wm);
        } else {
            // sampleCosineWeightedHemisphere, samplers.cuh:88-95
            float phi = 2.0f * KPT_PI * u.x;
            float ct = __builtin_sqrtf(u.y);
            float stt = __builtin_sqrtf(1.0f - u.y);
            v3 wi = mk3(stt * dm::cosf(phi), stt * dm::sinf(phi), ct);
            v3 f = b.albedo * f_diffuse(wo, wi);
            f = f * (1.0f / (1.0f - gp));
            bs.scatter = SC_DIFFUSE | SC_REFLECTED; bs.f = f; bs.wi = wi; bs.pdf = pdf_diffuse(wi);
        }
    }
    bs.wi = mul3(b.basis, bs.wi);
    return bs;
}
// UnifiedBSDF::regularize + GGX::regularize, bsdf.cuh:80-84, 271-276
KD void bsdf_regularize(BSDF& b) {
    if (b.roughness < 0.3f) b.roughness = clampf(2.0f * b.roughness, 0.1f, 0.3f);
    if (b.alpha < 0.3f) b.alpha = clampf(2.0f * b.alpha, 0.1f, 0.3f);
}

// ---------------------------------------------------------------- sky
// intersectSphere centred at the origin, atmosphere.cuh:30-57
KD bool isect_sphere(v3 o, v3 d, float radius, float* t0, float* t1) {
    v3 oc = mk3(0.0f, 0.0f, 0.0f) - o;
    float a = dot(d, d);
    float b = -2.0f * dot(d, oc);
    float c = dot(oc, oc) - sqr(radius);
    float disc = sqr(b) - 4.0f * a * c;
    if (disc < 0.0f) return false;
    float sq = __builtin_sqrtf(disc);
    float x0 = (-b - sq) / (2.0f * a);
    float x1 = (-b + sq) / (2.0f * a);
    if (x0 > x1) { float t = x0; x0 = x1; x1 = t; }
    *t0 = x0; *t1 = x1;
    if (x1 < 0.0f) return false;
    return true;
}
// NishitaAtmosphereModel::sampleLi, atmosphere.cuh:132-223
KD v3 sky_li(v3 o, v3 d, v3 sun, float emission) {
    const float eR = KPT_EARTH_R, aR = KPT_ATMOS_R, Hr = 7994.0f, Hm = 1200.0f;
    const v3 sR = mk3(3.8e-6f, 13.5e-6f, 33.1e-6f);
    const v3 sM = mk3(21e-6f, 21e-6f, 21e-6f);
    const float aRay = 0.0f, aMie = 21e-6f * 0.1f;
    float tmin = 0.0f, tmax = __builtin_inff();
    {
        float te = 0.0f, tx = 0.0f;
        if (!isect_sphere(o, d, aR, &te, &tx) || tx < 0.0f) return mk3(0.0f, 0.0f, 0.0f);
        if (te > tmin && te > 0.0f) tmin = te;
        if (tx < tmax) tmax = tx;
    }
    const float depth = tmax - tmin;
    const float step = depth / 16.0f;
    float cur = tmin;
    v3 iR = mk3(0.0f, 0.0f, 0.0f), iM = mk3(0.0f, 0.0f, 0.0f);
    float pRs = 0.0f, pMs = 0.0f;
    const v3 tR = sR + aRay;
    const v3 tM = sM + aMie;
    for (uint32_t i = 0; i < 16; ++i) {
        const v3 sp = o + ((cur + (step * 0.5f)) * d);
        const float alt = length(sp) - eR;
        const float pR = dm::expf(-alt / Hr) * step;
        const float pM = dm::expf(-alt / Hm) * step;
        pRs += pR;
        pMs += pM;
        float le = 0.0f, lx = 0.0f;
        isect_sphere(sp, sun, aR, &le, &lx);
        const float lstep = lx / 8.0f;
        float curl = 0.0f, lR = 0.0f, lM = 0.0f;
        uint32_t j;
        for (j = 0; j < 8; ++j) {
            const v3 spl = sp + ((curl + (lstep * 0.5f)) * sun);
            const float altl = length(spl) - eR;
            if (altl < 0.0f) break;
            lR += dm::expf(-altl / Hr) * lstep;
            lM += dm::expf(-altl / Hm) * lstep;
            curl += lstep;
        }
        if (j == 8) {
            const v3 tau = tR * (pRs + lR) + tM * (pMs + lM);
            const v3 Tr = mk3(dm::expf(-tau.x), dm::expf(-tau.y), dm::expf(-tau.z));
            iR = iR + Tr * pR;
            iM = iM + Tr * pM;
        }
        cur += step;
    }
    const float mu = dot(d, sun);
    const float ray_phase = (3.0f * (1.0f + sqr(mu))) / (16.0f * KPT_PI);
    const float g = 0.76f;
    const float denom = (8.0f * KPT_PI) * ((2.0f + sqr(g)) * dm::powf(1.0f + sqr(g) - 2.0f * g * mu, 1.5f));
    const float mie_phase = 3.0f * ((1.0f - sqr(g)) * (1.0f + sqr(mu))) / denom;
    const v3 L = (sR * iR * ray_phase + sM * iM * mie_phase);
    return L * KPT_SUN_NITS * emission;
}

// ---------------------------------------------------------------- surfaces
struct Surf {
    int32_t prim, inst, material, light;
    v3 wo, pos, n;
    v2 uv;
    float distance;
    bool backface;
};

// Intersection::getSurfaceInteraction, interaction.cu:11-38
KD Surf make_surf(const SceneView& S, const Hit& h, const Ray& r) {
    Surf s;
    const TriPos& tp = S.tris[h.slot];
    s.prim = tp.prim;
    s.inst = h.inst;
    s.distance = h.t;
    s.wo = -r.d;
    const TriAttr a = S.attrs[s.prim];
    const float* model = S.inst_shade[h.inst].model;
    s.material = a.material;
    s.light = a.light;
    s.pos = r.o + (h.t * r.d);
    s.n = normalize(mul4_xyz(model, mk4(a.lgn[0], a.lgn[1], a.lgn[2], 0.0f)));
    float b0 = 1.0f - h.u - h.v, b1 = h.u, b2 = h.v;
    s.uv.x = (b0 * a.uv0[0]) + (b1 * a.uv1[0]) + (b2 * a.uv2[0]);
    s.uv.y = (b0 * a.uv0[1]) + (b1 * a.uv1[1]) + (b2 * a.uv2[1]);
    s.backface = false;
    if (!(dot(s.n, s.wo) > 0.0f)) { s.n = s.n * -1.0f; s.backface = true; }
    return s;
}

// Texture::evaluate, texture.cuh:34-66
KD v3 tex_eval(const SceneView& S, int tex, v2 uv) {
    const Texture t = S.textures[tex];
    if (t.offset < 0 || t.offset >= S.num_pixel_bytes) return mk3(0.0f, 0.0f, 0.0f);
    int64_t pxs = (int64_t)t.channels * (t.bit_depth / 8);
    int64_t rows = (int64_t)t.width * pxs;
    float sx = uv.x - __builtin_floorf(uv.x);
    float sy = uv.y - __builtin_floorf(uv.y);
    int32_t x = int32_t(sx * (float)t.width) % t.width;
    int32_t y = int32_t(sy * (float)t.height) % t.height;
    if (x < 0) x += t.width;
    if (y < 0) y += t.height;
    int64_t off = (int64_t)y * rows + (int64_t)x * pxs;
    if (off + 3 >= S.num_pixel_bytes) return mk3(0.0f, 0.0f, 0.0f);
    const uint8_t* p = S.pixels + t.offset + off;
    return mk3((float)(int)p[0] * KPT_INV_255, (float)(int)p[1] * KPT_INV_255, (float)(int)p[2] * KPT_INV_255);
}
KD float srgb_decode(float v) {  // color.cuh:428-433
    return (v <= 0.04045f) ? (v / 12.92f) : dm::powf((v + 0.055f) / (1.0f + 0.055f), 2.4f);
}
// Material::getBSDF, material.cu:17-70
KD BSDF get_bsdf(const SceneView& S, const Surf& s) {
    const Material m = S.materials[s.material];
    v3 albedo = mk3(m.albedo[0], m.albedo[1], m.albedo[2]);
    float rough = m.roughness, metal = m.metallic, trans = m.transmission;
    v3 wg = s.n;
    if (m.albedo_tex >= 0) {
        v3 t = tex_eval(S, m.albedo_tex, s.uv);
        albedo = albedo * mk3(srgb_decode(t.x), srgb_decode(t.y), srgb_decode(t.z));
    }
    if (m.orm_tex >= 0) {
        v3 t = tex_eval(S, m.orm_tex, s.uv);
        rough *= t.y;
        metal *= t.z;
    }
    if (m.transmission_tex >= 0) {
        v3 t = tex_eval(S, m.transmission_tex, s.uv);
        trans *= t.x;
    }
    if (m.normal_tex >= 0) {
        v3 t = tex_eval(S, m.normal_tex, s.uv);
        v3 mn = (t * 2.0f) - 1.0f;
        mn.x *= m.normal_scale;
        mn.y *= m.normal_scale;
        m3 fr = onb_frisvad(wg);
        v3 sn = mul3(fr, normalize(mn));
        wg = (dot(s.wo, sn) > 0.0f) ? sn : wg;
    }
    BSDF b;
    b.basis = onb_frisvad(wg);
    b.albedo = albedo;
    b.metallic = metal;
    b.roughness = rough;
    b.transmission = trans;
    b.ior = m.ior;
    b.backface = s.backface;
    b.alpha = sqr(rough);
    return b;
}
// Light::L, light.cu:20-31
KD v3 light_L(const SceneView& S, const Light& l, v2 uv) {
    const Material m = S.materials[S.attrs[l.prim].material];
    v3 e = mk3(m.emission[0], m.emission[1], m.emission[2]) * m.emission_scale;
    if (m.emission_tex >= 0) {
        v3 t = tex_eval(S, m.emission_tex, uv);
        e = e * mk3(dm::powf(t.x, 2.2f), dm::powf(t.y, 2.2f), dm::powf(t.z, 2.2f));
    }
    return e;
}
// Light::pdf_Li, light.cu:11-18
KD float light_pdf_li(const Light& l, v3 ctx_pos, v3 lpos, v3 lnrm) {
    float area_pdf = 1.0f / l.area;
    float dist = distance(lpos, ctx_pos);
    float c = absdot(normalize(lpos - ctx_pos), lnrm);
    return area_pdf / (c / sqr(dist));
}
// clampOutput, color.cuh:295-302
KD v3 clamp_output(v3 v) {
    if (dm::f_isnan(v.x) || dm::f_isnan(v.y) || dm::f_isnan(v.z) ||
        dm::f_isinf(v.x) || dm::f_isinf(v.y) || dm::f_isinf(v.z)) return mk3(0.0f, 0.0f, 0.0f);
    return mk3(clampf(v.x, 0.0f, 1.0e8f), clampf(v.y, 0.0f, 1.0e8f), clampf(v.z, 0.0f, 1.0e8f));
}

// Camera::generateRay with getShadingJob's pixel-centre mapping
// (camera.cuh:16-26, shading_kernel.cuh:15-27).
KD Ray camera_ray_m(const float* inv_view, const float* inv_proj, const float* cam_pos, int width, int height,
                    int px, int py) {
    float u = ((float)px + 0.5f) / (float)width;
    float v = ((float)py + 0.5f) / (float)height;
    float nx = 2.0f * u - 1.0f, ny = 2.0f * v - 1.0f;
    v4 tcs = mul4(inv_proj, mk4(nx, ny, 1.0f, 1.0f));
    v4 tws = mul4(inv_view, mk4(normalize(xyz(tcs) / tcs.w), 0.0f));
    v3 dir = normalize(xyz(tws));
    return make_ray(mk3(cam_pos[0], cam_pos[1], cam_pos[2]), dir);
}
KD Ray camera_ray(const FrameConst& fc, int px, int py) {
    return camera_ray_m(fc.inv_view, fc.inv_proj, fc.cam_pos, fc.width, fc.height, px, py);
}

// GBuffer pack (packing.cuh:13-17, 105-125, 178-197)
KD uint32_t float_to_byte(float v) { v = saturate(v); return uint32_t(v * 255.0f) & 0xFFu; }
KD uint4 pack_gbuffer(v3 albedo, v3 normal, float depth) {
    float ax = __builtin_fabsf(normal.x), ay = __builtin_fabsf(normal.y), az = __builtin_fabsf(normal.z);
    float sum = ax + ay + az;
    float x = normal.x / sum, y = normal.y / sum;
    if (normal.z <= 0.0f) {
        float rx = 1.0f - __builtin_fabsf(y);
        float ry = 1.0f - __builtin_fabsf(x);
        x = x >= 0.0f ? rx : -rx;
        y = y >= 0.0f ? ry : -ry;
    }
    uint4 o;
    o.x = __builtin_bit_cast(uint32_t, x);
    o.y = __builtin_bit_cast(uint32_t, y);
    o.z = float_to_byte(albedo.x) | (float_to_byte(albedo.y) << 8) | (float_to_byte(albedo.z) << 16);
    o.w = __builtin_bit_cast(uint32_t, depth);
    return o;
}
// computeVelocity, kernels.cu:165-194 (camera products precomputed per frame)
KD v2 velocity(const SceneView& S, const FrameConst& fc, int inst, v3 pos) {
    v2 vel; vel.x = 0.0f; vel.y = 0.0f;
    if (inst < 0) return vel;
    const InstShade& is = S.inst_shade[inst];
    v3 local = mul4_xyz(is.curr_inv, mk4(pos, 1.0f));
    v4 cc = mul4(fc.curr_vp, mk4(pos, 1.0f));
    // (prev_vp * prev_model) * v: the matrix product first, as in the reference
    float pm[16];
    for (int c = 0; c < 4; ++c) {
        for (int r = 0; r < 4; ++r) {
            // Matrix4x4::operator*: element(r,c) = dot(row r of A, column c of B)
            pm[c * 4 + r] = fc.prev_vp[0 * 4 + r] * is.prev_model[c * 4 + 0] + fc.prev_vp[1 * 4 + r] * is.prev_model[c * 4 + 1] +
                            fc.prev_vp[2 * 4 + r] * is.prev_model[c * 4 + 2] + fc.prev_vp[3 * 4 + r] * is.prev_model[c * 4 + 3];
        }
    }
    v4 pc = mul4(pm, mk4(local, 1.0f));
    v3 cn = xyz(cc) / cc.w, pn = xyz(pc) / pc.w;
    float cux = (cn.x + 1.0f) / 2.0f, cuy = (cn.y + 1.0f) / 2.0f;
    float pux = (pn.x + 1.0f) / 2.0f, puy = (pn.y + 1.0f) / 2.0f;
    vel.x = cux - pux; vel.y = cuy - puy;
    return vel;
}

}  // namespace kpt
