// kpt_integrator.h — the per-path integrator (Integrator::Li,
// integrator.cuh:241-402) as device code, split into the pieces the
// megakernel and the wavefront stages share.
#pragma once
#include "kpt_cache.h"
#include "kpt_shade.h"

namespace kpt {

struct VisOut {
    v3 albedo;
    v3 normal;
    v3 pos;
    float depth;
    int32_t inst;
    int32_t prim;
};

KD v3 observer_pos() { return mk3(0.0f, KPT_EARTH_R + 1.0f, 0.0f); }

// Sky radiance for a miss, including LeSun on the primary ray
// (integrator.cuh:315-326, 135-143).
KD v3 miss_radiance(const FrameConst& fc, v3 dir, bool first) {
    const v3 sun = ld3(fc.sun_dir);
    v3 sky = sky_li(observer_pos(), dir, sun, fc.sun_emission);
    if (first) {
        float sim = dot(dir, sun);
        bool on_sun = sim > fc.min_sun_cos;
        float sr = fc.sun_emission * KPT_SUN_NITS;
        float le = sr * dm::powf(1.0f / sr, on_sun ? 0.0f : 1.0f);
        sky = sky * mk3(le, le, le);
    }
    return sky;
}

// A deferred next-event estimate: the shadow ray and the two possible
// additions to L (unoccluded / occluded).  When no shadow ray is needed,
// `trace` is false and `occl` is the addition.
struct NeeResult {
    bool trace;
    v3 org;
    v3 dir;
    float tmax;
    v3 unoccl;
    v3 occl;
};

// sampleLdSun / sampleLd and the L += Ld*beta/p of integrator.cuh:365-375,
// with the visibility test left to the caller.  Consumes the same random
// numbers, in the same order, as the reference.  `sun_p` is the probability
// of sampling the sun rather than an area light: the reference's 0.5 on every
// path vertex; the radiance cache's vertex passes 1 when the scene has no area
// lights (same random numbers: the selection draw is < 1 always).
KD NeeResult next_event(const SceneView& S, const FrameConst& fc, const Surf& s, const BSDF& b, v3 ray_d, v3 beta,
                        uint32_t& seed, float sun_p = 0.5f) {
    NeeResult r;
    r.trace = false;
    const v3 zero = mk3(0.0f, 0.0f, 0.0f);
    if (rng1(seed) < sun_p) {
        // sampleLdSun, integrator.cuh:147-186
        const v3 sun = ld3(fc.sun_dir);
        v3 sun_pos = sun * KPT_SOLAR_DIST;
        float sun_radius = fc.sun_radius;
        float ux = rng1(seed);
        float uy = rng1(seed);
        float uz = rng1(seed);
        v3 off = mk3(ux * 2.0f - 1.0f, uy * 2.0f - 1.0f, uz * 2.0f - 1.0f);
        v3 target = sun_pos + (off * sun_radius);
        v3 wo = -ray_d;
        v3 fcos = bsdf_f(b, wo, sun) * fmaxf_(0.0f, dot(sun, (s.backface ? -1.0f : 1.0f) * s.n));
        r.occl = zero * beta / sun_p;
        if (is_zero(fcos)) return r;
        v3 srad = ld3(fc.sun_radiance);
        if (is_zero(srad)) return r;
        // spawnRayTo + Unoccluded (interaction.cuh:43-53, integrator.cuh:124-131)
        v3 org = s.backface ? s.pos - (s.n * KPT_HIT_EPS) : s.pos + (s.n * KPT_HIT_EPS);
        r.org = org;
        r.dir = normalize(target - org);
        r.tmax = distance(target, org) - 0.11f;
        float sun_area = KPT_PI * sqr(sun_radius);
        v3 snrm = normalize(target - sun_pos), wi = normalize(target - s.pos);
        float cth = absdot(snrm, -wi);
        float pdf = (1.0f / sun_area) / (cth / sqr(KPT_SOLAR_DIST));
        v3 Ld = (fcos * srad * KPT_SUN_NITS) / pdf;
        r.unoccl = Ld * beta / sun_p;
        r.trace = true;
        return r;
    }
    // sampleLd, integrator.cuh:188-220
    const float p_other = 1.0f - sun_p;
    r.occl = zero * beta / p_other;
    float ul = rng1(seed);
    const int nl = fc.num_lights;
    if (nl < 1) return r;
    int li = (int)(ul * (float)nl);
    if (li > nl - 1) li = nl - 1;
    const Light l = S.lights[li];
    float pmf = 1.0f / (float)nl;
    float u2x = rng1(seed);
    float u2y = rng1(seed);
    // Triangle::sample (triangle.cuh:104-119) through Light::sampleLi (light.cu:33-45)
    const TriAttr ta = S.attrs[l.prim];
    const float* model = S.inst_shade[l.inst].model;
    float bx, by;
    if (u2x < u2y) { bx = u2x / 2.0f; by = u2y - bx; } else { by = u2y / 2.0f; bx = u2x - by; }
    float bz = 1.0f - bx - by;
    v2 luv;
    luv.x = (bx * ta.uv0[0]) + (by * ta.uv1[0]) + (bz * ta.uv2[0]);
    luv.y = (bx * ta.uv0[1]) + (by * ta.uv1[1]) + (bz * ta.uv2[1]);
    // object-space vertices of the light triangle (v0 + e1 need not equal v1
    // in float, so lights keep their own vertex copy)
    v3 p0 = ld3(S.light_verts + 9 * (size_t)li), p1 = ld3(S.light_verts + 9 * (size_t)li + 3),
       p2 = ld3(S.light_verts + 9 * (size_t)li + 6);
    v3 p = (p0 * bx) + (p1 * by) + (p2 * bz);
    p = mul4_xyz(model, mk4(p, 1.0f));
    v3 gn = mul4_xyz(model, mk4(ta.lgn[0], ta.lgn[1], ta.lgn[2], 0.0f));
    float area = 0.5f * length(cross(p1 - p0, p2 - p0));
    float spdf = 1.0f / area;
    spdf *= sqr(distance(p, s.pos));
    spdf /= absdot(normalize(p - s.pos), gn);
    v3 wi = normalize(p - s.pos);
    v3 lL = light_L(S, l, luv);
    if (spdf <= 0.0f || is_zero(lL)) return r;
    v3 wo = -ray_d;
    v3 fcos = bsdf_f(b, wo, wi) * fmaxf_(0.0f, dot(wi, (s.backface ? -1.0f : 1.0f) * s.n));
    if (is_zero(fcos)) return r;
    v3 org = s.backface ? s.pos - (s.n * KPT_HIT_EPS) : s.pos + (s.n * KPT_HIT_EPS);
    r.org = org;
    r.dir = normalize(p - org);
    r.tmax = distance(p, org) - 0.11f;
    float pl = pmf * spdf;
    float pb = bsdf_pdf(b, wo, wi);
    float f = 1.0f * pl, g = 1.0f * pb;
    float wl = sqr(f) / (sqr(f) + sqr(g));
    v3 Ld = (lL * fcos * wl) / pl;
    r.unoccl = Ld * beta / p_other;
    r.trace = true;
    return r;
}

// Emission of a hit with MIS against the previous BSDF sample
// (integrator.cuh:333-343).  Returns the addition to L (zero when none).
KD bool emitted(const SceneView& S, const FrameConst& fc, const Surf& s, bool first, float p_b, v3 prev_pos,
                v3 beta, v3* Le_out, v3* add) {
    v3 Le = mk3(0.0f, 0.0f, 0.0f);
    if (s.light >= 0) Le = light_L(S, S.lights[s.light], s.uv);
    *Le_out = Le;
    if (is_zero(Le)) return false;
    float wl = 1.0f;
    if (s.light >= 0 && !first) {
        float lpdf = (1.0f / (float)fc.num_lights) * light_pdf_li(S.lights[s.light], prev_pos, s.pos, s.n);
        float f = 1.0f * p_b, g = 1.0f * lpdf;
        wl = sqr(f) / (sqr(f) + sqr(g));
    }
    *add = Le * wl * beta;
    return true;
}

// Li's loop state between vertices: what the wavefront keeps in HBM.
struct PathState {
    Ray ray;
    v3 L, beta;
    bool any_nonspec;
    int32_t depth;
    float p_b;
    v3 prev_pos;
    bool wide_in = false;   // radiance cache: this vertex was reached through a wide lobe
    bool cached = false;    // radiance cache: the path has made its training record
};

// The radiance cache at a path vertex (no reference counterpart).  A path's
// cache vertex is the first vertex after the primary hit reached through a
// wide lobe (a diffuse sample, or a lobe of roughness >= cf.wide_roughness),
// where the lobe blurs the network's errors.  There a training path (chosen
// by a hash of path and frame against the frame's threshold, which
// k_rc_frame_start leaves in cb.counters[2]) records the
// vertex and continues: its suffix becomes the network's target.  Any other
// path, when the cache renders, writes a query and ends there: the network
// completes it after the frame.  A training path ends at its next cache
// vertex in the network (self-training).  The network learns radiance
// divided by the throughput and a reflectance proxy of the vertex
// (0.04 + 0.96 albedo).  With cf.nee the vertex's next-event estimate is
// traced exactly first and the network stands for the continuation only; the
// estimate then samples only the sun when the scene has no area lights (the
// reference's 50 % area-light draw would be a wasted, zero sample).
struct CacheVertex {
    bool to_cache;    // the path ends here in the network (its query is written)
    int32_t trec;     // training record made here (-1: none)
    float sun_p;      // sun probability of this vertex's next-event estimate
};
KD CacheVertex cache_vertex(const FrameConst& fc, const CacheBuffers& cb, const CacheFrame& cf, int32_t p,
                            int32_t depth, bool wide_in, bool& cached, const Surf& s, const BSDF& b, v3 ray_d,
                            v3 beta) {
    CacheVertex cv;
    cv.to_cache = false;
    cv.trec = -1;
    if (depth >= 1 && wide_in && depth < fc.max_depth && (cf.enabled || cf.train) && !(cached && !cf.enabled)) {
        const uint32_t hs = pcg_hash((uint32_t)p ^ (cf.frame * 0x9E3779B9u)) & 0xFFFFFFu;
        const float4 qpos = make_float4(s.pos.x, s.pos.y, s.pos.z, b.roughness);
        const float4 qdir = make_float4(-ray_d.x, -ray_d.y, -ray_d.z, __builtin_bit_cast(float, (uint32_t)p));
        const v3 bq = beta * (mk3(0.04f, 0.04f, 0.04f) + b.albedo * 0.96f);
        if (!cached && cf.train && hs < (uint32_t)cb.counters[2]) {
            cached = true;
            const int r = atomicAdd(&cb.counters[1], 1);
            if (r < cb.train_cap) {
                cb.t_pos[r] = qpos;
                cb.t_dir[r] = qdir;
                cb.t_beta[r] = make_float4(bq.x, bq.y, bq.z, 0.0f);
                cv.trec = r;
            }
        } else if (cf.enabled) {
            const int r = atomicAdd(&cb.counters[0], 1);
            if (r < cb.query_cap) {
                cb.q_pos[r] = qpos;
                cb.q_dir[r] = qdir;
                cb.q_beta[r] = make_float4(bq.x, bq.y, bq.z, 0.0f);
                cv.to_cache = true;
            }
        }
    }
    cv.sun_p = (cf.nee && (cv.to_cache || cv.trec >= 0) && fc.num_lights == 0) ? 1.0f : 0.5f;
    return cv;
}

// One vertex of Li's loop (integrator.cuh:266-399): trace from the path's
// ray, add emission and the next-event estimate, sample the next direction.
// Returns false where the reference's loop breaks (miss, last depth, zero
// BSDF sample, Russian roulette); the caller also stops when beta is zero.
// DEFER_SKY (vertices past the first only): a miss sets *missed and leaves
// L without the sky term — the last addition of the path, L + sky * beta,
// which the caller then makes with ps.ray.d and ps.beta as they are.
// CACHE: the radiance cache's hooks at the path's cache vertex (cache_vertex
// with path index `path`), as the wavefront's shade stage applies them; a
// path that ends in the network returns false after its next-event estimate.
// heat (COUNT with cnt): the path's pixel's DebugData pair, to which the
// closest-hit trace's BLAS and TLAS node pops are added.
// li_shade is the vertex after its closest-hit trace (h, from ps.ray): the
// part li_bounce and the wavefront tail share.
template <bool COUNT = false, bool DEFER_SKY = false, bool CACHE = false>
KD bool li_shade(const SceneView& S, const FrameConst& fc, PathState& ps, uint32_t& seed, const Hit& h,
                 const StackShape& lds, bool* ovf, VisOut* vs, TraceCounts* cnt = nullptr, bool* missed = nullptr,
                 const CacheBuffers* cb = nullptr, const CacheFrame* cf = nullptr, int32_t path = -1,
                 uint32_t* spill = nullptr);

template <bool COUNT = false, bool DEFER_SKY = false, bool CACHE = false>
KD bool li_bounce(const SceneView& S, const FrameConst& fc, PathState& ps, uint32_t& seed, const StackShape& lds,
                  bool* ovf, VisOut* vs, TraceCounts* cnt = nullptr, bool* missed = nullptr,
                  const CacheBuffers* cb = nullptr, const CacheFrame* cf = nullptr, int32_t path = -1,
                  int32_t* heat = nullptr) {
    seed = seed + (uint32_t)ps.depth;
    bool o = false;
    const uint32_t b0 = (COUNT && cnt) ? cnt->interior + cnt->leaf : 0u, t0 = (COUNT && cnt) ? cnt->tlas : 0u;
    Hit h = trace_closest<COUNT>(S, ps.ray, 0.0f, __builtin_inff(), lds, &o, cnt);
    if (COUNT && cnt && heat) {
        atomicAdd(&heat[0], (int32_t)(cnt->interior + cnt->leaf - b0));
        atomicAdd(&heat[1], (int32_t)(cnt->tlas - t0));
    }
    *ovf = *ovf || o;
    return li_shade<COUNT, DEFER_SKY, CACHE>(S, fc, ps, seed, h, lds, ovf, vs, cnt, missed, cb, cf, path);
}

template <bool COUNT, bool DEFER_SKY, bool CACHE>
KD bool li_shade(const SceneView& S, const FrameConst& fc, PathState& ps, uint32_t& seed, const Hit& h,
                 const StackShape& lds, bool* ovf, VisOut* vs, TraceCounts* cnt, bool* missed,
                 const CacheBuffers* cb, const CacheFrame* cf, int32_t path, uint32_t* spill) {
    const float eta_scale = 1.0f;
    const bool first = (ps.depth == 0);
    if (DEFER_SKY && h.slot < 0 && !first) {
        *missed = true;
        return false;
    }
    if (h.slot < 0) {
        v3 sky = miss_radiance(fc, ps.ray.d, first);
        if (first) {
            vs->albedo = sky; vs->depth = __builtin_inff(); vs->normal = mk3(0.0f, 0.0f, 0.0f);
            vs->pos = mk3(0.0f, 0.0f, 0.0f); vs->inst = -1; vs->prim = -1;
        }
        ps.L = ps.L + sky * ps.beta;
        return false;
    }
    Surf s = make_surf(S, h, ps.ray);
    v3 Le, add;
    if (emitted(S, fc, s, first, ps.p_b, ps.prev_pos, ps.beta, &Le, &add)) ps.L = ps.L + add;
    BSDF b = get_bsdf(S, s);
    if (first) {
        vs->albedo = !is_zero(Le) ? Le : b.albedo;
        vs->depth = s.distance; vs->normal = s.n; vs->pos = s.pos; vs->inst = s.inst; vs->prim = s.prim;
        b.albedo = mk3(1.0f, 1.0f, 1.0f);
    }
    if (ps.any_nonspec) bsdf_regularize(b);
    CacheVertex cv;
    cv.to_cache = false;
    cv.trec = -1;
    cv.sun_p = 0.5f;
    if (CACHE) {
        cv = cache_vertex(fc, *cb, *cf, path, ps.depth, ps.wide_in, ps.cached, s, b, ps.ray.d, ps.beta);
        if (cv.trec >= 0 && !cf->nee) {
            cb->t_lbefore[cv.trec] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
            cv.trec = -1;
        }
    }
    // The next-event estimate's random numbers are drawn here, before the
    // continuation's, as the reference draws them; its shadow ray is traced
    // after the continuation is sampled (its answer only picks the term added
    // to L, and nothing else touches L in between), so the surface and BSDF
    // are no longer live while it runs.
    NeeResult ne;
    ne.trace = false;
    bool nee = false;
    if (!(CACHE && cv.to_cache && !cf->nee) && !(b.alpha < 0.001f)) {
        ne = next_event(S, fc, s, b, ps.ray.d, ps.beta, seed, cv.sun_p);
        nee = true;
    }
    bool more = true;
    if (ps.depth++ == fc.max_depth) more = false;
    else if (CACHE && cv.to_cache) more = false;
    else {
        v3 wo = -ps.ray.d;
        v2 u2, X2;
        u2.x = rng1(seed); u2.y = rng1(seed);
        X2.x = rng1(seed); X2.y = rng1(seed);
        BSample bs = bsdf_sample(b, wo, u2, X2);
        if (is_zero(bs.f)) more = false;
        else {
            ps.beta = ps.beta * (bs.f * absdot(s.n, bs.wi) / bs.pdf);
            ps.p_b = bsdf_pdf(b, wo, bs.wi);
            ps.any_nonspec = ps.any_nonspec || !((bs.scatter & SC_SPECULAR) != 0);
            if (CACHE) ps.wide_in = ((bs.scatter & SC_DIFFUSE) != 0) || b.roughness >= cf->wide_roughness;
            ps.prev_pos = s.pos;
            v3 org = (bs.scatter & SC_TRANSMITTED) ? s.pos - (s.n * KPT_HIT_EPS) : s.pos + (s.n * KPT_HIT_EPS);
            ps.ray = make_ray(org, bs.wi);
            v3 rr = ps.beta * eta_scale;
            if (max_comp(rr) < 1.0f && ps.depth > 1) {
                float q = fmaxf_(0.0f, 1.0f - max_comp(rr));
                if (rng1(seed) < q) more = false;
                else ps.beta = ps.beta / (1.0f - q);
            }
        }
    }
    if (nee) {
        if (ne.trace) {
            bool oo = false;
            Ray sr = make_ray(ne.org, ne.dir);
            // the caller's stack spill area when it has one (its closest-hit trace has ended)
            bool occluded = spill ? trace_any_in<COUNT>(S, sr, 0.0f, ne.tmax, lds, spill, &oo, cnt)
                                  : trace_any<COUNT>(S, sr, 0.0f, ne.tmax, lds, &oo, cnt);
            *ovf = *ovf || oo;
            ps.L = ps.L + (occluded ? ne.occl : ne.unoccl);
        } else {
            ps.L = ps.L + ne.occl;
        }
    }
    // a training record's radiance so far includes the vertex's estimate
    if (CACHE && cv.trec >= 0) cb->t_lbefore[cv.trec] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
    return more;
}

// Li from a path state to the end of the path, in one thread (the megakernel
// from depth 0).  ps.L holds the radiance on return.
KD v3 li_loop(const SceneView& S, const FrameConst& fc, PathState& ps, uint32_t& seed, const StackShape& lds, bool* ovf,
              VisOut* vs) {
    while (!is_zero(ps.beta)) {
        if (!li_bounce(S, fc, ps, seed, lds, ovf, vs)) break;
    }
    return ps.L;
}

// Megakernel form of Li: the whole path in one thread.
KD v3 li_path(const SceneView& S, const FrameConst& fc, Ray ray, uint32_t& seed, const StackShape& lds, bool* ovf,
              VisOut* vs) {
    PathState ps;
    ps.ray = ray;
    ps.L = mk3(0.0f, 0.0f, 0.0f);
    ps.beta = mk3(1.0f, 1.0f, 1.0f);
    ps.any_nonspec = false;
    ps.depth = 0;
    ps.p_b = 1.0f;
    ps.prev_pos = mk3(0.0f, 0.0f, 0.0f);
    return li_loop(S, fc, ps, seed, lds, ovf, vs);
}

}  // namespace kpt
