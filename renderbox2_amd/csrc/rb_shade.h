// rb_shade.h -- device-side shading helpers of the render passes: triangle
// fetch, area-light pdf, Lambertian / mirror BSDFs (the subset the
// configurations use), in fp32 with every rounding written out.
#pragma once
#include "rb_device.h"
#include "rb_sample.h"

namespace rb {

struct SceneDev {
    const float4* __restrict__ tri_v;  // 3 float4 per triangle (v0|material, v1, v2)
    const rb_material* __restrict__ mats;
    const uint32_t* __restrict__ light_tris;
    uint32_t n_lights;
};

__device__ __forceinline__ void tri_fetch(const SceneDev& S, uint32_t t, V3* v0, V3* v1, V3* v2, uint32_t* mat) {
    const float4 a = S.tri_v[3 * (size_t)t], b = S.tri_v[3 * (size_t)t + 1], c = S.tri_v[3 * (size_t)t + 2];
    *v0 = v3(a.x, a.y, a.z);
    *v1 = v3(b.x, b.y, b.z);
    *v2 = v3(c.x, c.y, c.z);
    if (mat) *mat = __float_as_uint(a.w);
}

RB_HD float vlen(V3 a) { return sqrtf(vdot(a, a)); }
RB_HD bool is_black(V3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
// luminance (src/core/util.h:161-169)
RB_HD float luminance(V3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

// sample_triangle_pdf (src/core/montecarlo.cu:166-182): solid-angle pdf of
// direction `d` from `p` towards the triangle: |p - hit|^2 / (area * cos).
// Moller-Trumbore as intersect_ray_triangle (src/core/primitives.cu:51-70),
// here also requiring t > 0, and the triangle's true area (the reference's
// compute_area takes the length of a normalised cross product, always 0.5).
RB_HD float light_pdf(V3 p, V3 d, V3 v0, V3 v1, V3 v2, V3 ln, float area) {
    const V3 e1 = vsub(v1, v0), e2 = vsub(v2, v0);
    const V3 pv = vcross(d, e2);
    const float det = vdot(e1, pv);
    if (det < 1.0e-6f && det > -1.0e-6f) return 0.0f;
    const float inv = 1.0f / det;
    const V3 tv = vsub(p, v0);
    const float u = vdot(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return 0.0f;
    const V3 qv = vcross(tv, e1);
    const float v = vdot(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0.0f;
    const float t = vdot(e2, qv) * inv;
    if (!(t > 0.0f)) return 0.0f;
    const V3 h = vadd(p, vmul(d, t));
    const V3 dd = vsub(h, p);
    const float r2 = vdot(dd, dd);
    const float cth = fabsf(vdot(ln, vneg(d)));
    return r2 / (area * cth);
}

RB_HD float tri_area(V3 v0, V3 v1, V3 v2) { return 0.5f * vlen(vcross(vsub(v1, v0), vsub(v2, v0))); }

// evaluate_lambertian (src/core/reflection.cu:42-48) with the material_f
// reflection test on the geometric normal.
RB_HD V3 lambert_f(V3 wo, V3 wi, V3 ng, const rb_material& m) {
    if (!(vdot(wo, ng) * vdot(wi, ng) > 0.0f)) return v3(0.0f, 0.0f, 0.0f);
    const float k = 0.318309886183790672f;  // 1/pi
    return v3(m.color[0] * k, m.color[1] * k, m.color[2] * k);
}

// ------------------------------------------------------------ BSDFs ----
// Non-delta BSDFs (Lambert, Oren-Nayar, Blinn microfacet) reflect only and
// are evaluated in a frame about the normal facing wo (nf); mirror and glass
// are delta BSDFs, sampled only.  Reference: src/core/reflection.cu:42-300,
// dispatch src/core/material.cu:56-326 (single-layer materials).

RB_HD bool is_delta(const rb_material& m) { return m.kind == RB_MAT_MIRROR || m.kind == RB_MAT_GLASS; }

struct Frame {
    V3 b1, b2, n;
};
RB_HD Frame frame_of(V3 n) {
    Frame f;
    f.n = n;
    onb(n, &f.b1, &f.b2);
    return f;
}
RB_HD V3 to_local(const Frame& f, V3 v) { return v3(vdot(v, f.b1), vdot(v, f.b2), vdot(v, f.n)); }
RB_HD V3 to_world(const Frame& f, V3 l) { return vadd(vadd(vmul(f.b1, l.x), vmul(f.b2, l.y)), vmul(f.n, l.z)); }

// sin_theta / cos_phi / sin_phi of a local direction (src/core/reflection.h:78-100)
RB_HD float sin_theta_l(V3 w) {
    const float s2 = 1.0f - w.z * w.z;
    return sqrtf(s2 > 0.0f ? s2 : 0.0f);
}
RB_HD float cos_phi_l(V3 w) {
    const float st = sin_theta_l(w);
    if (st == 0.0f) return 1.0f;
    const float c = w.x / st;
    return c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
}
RB_HD float sin_phi_l(V3 w) {
    const float st = sin_theta_l(w);
    if (st == 0.0f) return 0.0f;
    const float c = w.y / st;
    return c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
}

// evaluate_orennayar (reflection.cu:144-178); A = param[1], B = param[2].
RB_HD V3 orennayar_local(V3 wi, V3 wo, const rb_material& m) {
    const float sti = sin_theta_l(wi), sto = sin_theta_l(wo);
    float maxcos = 0.0f;
    if (sti > 1.0e-4f && sto > 1.0e-4f) {
        const float dcos = cos_phi_l(wi) * cos_phi_l(wo) + sin_phi_l(wi) * sin_phi_l(wo);
        maxcos = dcos > 0.0f ? dcos : 0.0f;
    }
    float sinalpha, tanbeta;
    if (fabsf(wi.z) > fabsf(wo.z)) {
        sinalpha = sto;
        tanbeta = sti / fabsf(wi.z);
    } else {
        sinalpha = sti;
        tanbeta = sto / fabsf(wo.z);
    }
    const float k = m.param[1] + m.param[2] * maxcos * sinalpha * tanbeta;
    const float ip = 0.318309886183790672f;
    return v3(m.color[0] * ip * k, m.color[1] * ip * k, m.color[2] * ip * k);
}

// Fresnel::evaluate for a microfacet layer (src/core/reflection.h:147-217).
RB_HD V3 fresnel_layer(const rb_material& m, float cosi) {
    const float type = m.param[1];
    if (type == (float)RB_FRESNEL_CONDUCTOR) {
        float rc[3];
        for (int c = 0; c < 3; c++) {
            const float eta = m.param[2 + c], k = m.param[5 + c];
            const float temp = (eta * eta + k * k) * cosi * cosi;
            const float rparl2 = (temp - 2.0f * eta * cosi + 1.0f) / (temp + 2.0f * eta * cosi + 1.0f);
            const float tf = eta * eta + k * k;
            const float rperp2 = (tf - 2.0f * eta * cosi + cosi * cosi) / (tf + 2.0f * eta * cosi + cosi * cosi);
            rc[c] = (rparl2 + rperp2) / 2.0f;
        }
        return v3(rc[0], rc[1], rc[2]);
    }
    if (type == (float)RB_FRESNEL_DIELECTRIC) {
        float ei = m.param[2], et = m.param[3];
        if (!(cosi > 0.0f)) {
            const float t = ei;
            ei = et;
            et = t;
        }
        const float c2 = 1.0f - cosi * cosi;
        const float sint = ei / et * sqrtf(c2 > 0.0f ? c2 : 0.0f);
        if (sint >= 1.0f) return v3(1.0f, 1.0f, 1.0f);
        const float t2 = 1.0f - sint * sint;
        const float cost = sqrtf(t2 > 0.0f ? t2 : 0.0f);
        const float ci = fabsf(cosi);
        const float rparl = ((et * ci) - (ei * cost)) / ((et * ci) + (ei * cost));
        const float rperp = ((ei * ci) - (et * cost)) / ((ei * ci) + (et * cost));
        const float f = (rparl * rparl + rperp * rperp) / 2.0f;
        return v3(f, f, f);
    }
    return v3(1.0f, 1.0f, 1.0f);
}

// Blinn distribution (reflection.cu:198-240) and Torrance-Sparrow shadowing
// (:243-251, with the cos(theta_o) factor the reference drops from the first
// term); float 2*pi*4 in the pdf (the reference divides in double).
RB_HD float blinn_d(V3 wh, float e) { return (e + 2.0f) * 0.159154943091895336f * pow_series(fabsf(wh.z), e); }
RB_HD float micro_g(V3 wi, V3 wo, V3 wh) {
    const float nh = fabsf(wh.z), no = fabsf(wo.z), ni = fabsf(wi.z), oh = fabsf(vdot(wo, wh));
    const float g1 = 2.0f * nh * no / oh, g2 = 2.0f * nh * ni / oh;
    const float g = g1 < g2 ? g1 : g2;
    return g < 1.0f ? g : 1.0f;
}
RB_HD V3 microfacet_local(V3 wi, V3 wo, const rb_material& m) {
    const float ci = fabsf(wi.z), co = fabsf(wo.z);
    if (ci == 0.0f || co == 0.0f) return v3(0.0f, 0.0f, 0.0f);
    V3 wh = vadd(wi, wo);
    if (wh.x == 0.0f && wh.y == 0.0f && wh.z == 0.0f) return v3(0.0f, 0.0f, 0.0f);
    wh = vnorm(wh);
    const V3 F = fresnel_layer(m, vdot(wi, wh));
    const float D = blinn_d(wh, m.param[0]), G = micro_g(wi, wo, wh);
    const float den = 4.0f * ci * co;
    return v3(m.color[0] * D * G * F.x / den, m.color[1] * D * G * F.y / den, m.color[2] * D * G * F.z / den);
}
RB_HD float microfacet_pdf_local(V3 wi, V3 wo, float e) {
    if (!(wi.z * wo.z > 0.0f)) return 0.0f;
    const V3 wh = vnorm(vadd(wo, wi));
    const float d = vdot(wo, wh);
    if (d <= 0.0f) return 0.0f;
    return ((e + 1.0f) * pow_series(fabsf(wh.z), e)) / (6.28318531f * 4.0f * d);
}

// BSDF value for wo -> wi (world), with the reflection test on ng.
RB_HD V3 bsdf_f(const rb_material& m, V3 wo, V3 wi, V3 ng, V3 nf) {
    if (m.kind == RB_MAT_LAMBERT) return lambert_f(wo, wi, ng, m);
    if (is_delta(m) || m.kind == RB_MAT_EMITTER) return v3(0.0f, 0.0f, 0.0f);
    if (!(vdot(wo, ng) * vdot(wi, ng) > 0.0f)) return v3(0.0f, 0.0f, 0.0f);
    const Frame fr = frame_of(nf);
    const V3 wil = to_local(fr, wi), wol = to_local(fr, wo);
    return m.kind == RB_MAT_ORENNAYAR ? orennayar_local(wil, wol, m) : microfacet_local(wil, wol, m);
}

// Solid-angle pdf of sampling wi from wo with bsdf_sample.
RB_HD float bsdf_pdf(const rb_material& m, V3 wo, V3 wi, V3 nf) {
    if (m.kind == RB_MAT_MICROFACET) {
        const Frame fr = frame_of(nf);
        return microfacet_pdf_local(to_local(fr, wi), to_local(fr, wo), m.param[0]);
    }
    return fabsf(vdot(wi, nf)) * 0.318309886183790672f;
}

// Samples wi for a non-delta material; f and pdf as bsdf_f / bsdf_pdf.
RB_HD void bsdf_sample(const rb_material& m, V3 wo, V3 ng, V3 nf, float u1, float u2, V3* wi, V3* f, float* pdf) {
    if (m.kind != RB_MAT_MICROFACET) {  // cosine-weighted (sample_lambertian / sample_orennayar)
        *wi = cosine_hemisphere(nf, u1, u2);
        *f = bsdf_f(m, wo, *wi, ng, nf);
        *pdf = fabsf(vdot(*wi, nf)) * 0.318309886183790672f;
        return;
    }
    // sample_blinn_microfacet (reflection.cu:208-225)
    const Frame fr = frame_of(nf);
    const V3 wol = to_local(fr, wo);
    const float e = m.param[0];
    const float ct = pow_series(u1, 1.0f / (e + 1.0f));
    const float s2 = 1.0f - ct * ct;
    const float st = sqrtf(s2 > 0.0f ? s2 : 0.0f);
    float sp, cp;
    sincos_2pi(u2, &sp, &cp);
    V3 wh = v3(st * cp, st * sp, ct);
    if (!(wol.z * wh.z > 0.0f)) wh = vneg(wh);
    const float d = vdot(wol, wh);
    const V3 wil = vadd(vneg(wol), vmul(wh, 2.0f * d));
    *pdf = d > 0.0f ? ((e + 1.0f) * pow_series(ct, e)) / (6.28318531f * 4.0f * d) : 0.0f;
    *wi = vnorm(to_world(fr, wil));
    *f = (wil.z * wol.z > 0.0f) ? bsdf_f(m, wo, *wi, ng, nf) : v3(0.0f, 0.0f, 0.0f);
}

// Glass (sample_glass, reflection.cu:109-140, with the Fresnel-weighted
// choice between reflection and refraction that the reference omits): the
// outward normal ng orients the interface; returns the throughput factor
// f |cos| / pdf.  Radiance transport (camera paths) scales refraction by
// (eta_wo / eta_wi)^2, importance transport (light paths) does not.
RB_HD V3 glass_sample(const rb_material& m, V3 wo, V3 ng, float u, bool radiance, V3* wi) {
    const Frame fr = frame_of(ng);
    const V3 wol = to_local(fr, wo);
    const float cosi = wol.z;
    const bool entering = cosi > 0.0f;
    const float ei = entering ? m.param[0] : m.param[1], et = entering ? m.param[1] : m.param[0];
    rb_material fm = m;
    fm.param[1] = (float)RB_FRESNEL_DIELECTRIC;
    fm.param[2] = m.param[0];
    fm.param[3] = m.param[1];
    const float F = fresnel_layer(fm, cosi).x;
    const float s2 = 1.0f - cosi * cosi;
    const float eta = ei / et;
    const float st2 = eta * eta * (s2 > 0.0f ? s2 : 0.0f);
    V3 wil;
    V3 w;
    if (u < F || st2 >= 1.0f) {
        wil = v3(-wol.x, -wol.y, wol.z);
        w = v3(m.color[0], m.color[1], m.color[2]);
    } else {
        const float c2 = 1.0f - st2;
        float cost = sqrtf(c2 > 0.0f ? c2 : 0.0f);
        if (entering) cost = -cost;
        wil = v3(eta * -wol.x, eta * -wol.y, cost);
        const float sc = radiance ? (ei * ei) / (et * et) : 1.0f;
        w = v3(m.color[0] * sc, m.color[1] * sc, m.color[2] * sc);
    }
    *wi = vnorm(to_world(fr, wil));
    return w;
}

}  // namespace rb
