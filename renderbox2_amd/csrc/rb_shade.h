// rb_shade.h -- device-side shading helpers of the render passes: triangle
// fetch, area-light pdf, Lambertian / mirror BSDFs (the subset the
// configurations use), in fp32 with every rounding written out.
#pragma once
#include "rb_device.h"
#include "rb_sample.h"

namespace rb {

struct SceneDev {
    const float4* __restrict__ tri_v;  // 3 float4 per triangle (v0|material, v1|flags, v2)
    const rb_material* __restrict__ mats;
    const uint32_t* __restrict__ light_tris;
    uint32_t n_lights;
    // 3 float4 per triangle (n0, n1, n2 | 0) when the scene has shading
    // normals, else nullptr; a triangle has them when its kTriHasNormals flag is set
    const float4* __restrict__ tri_n;
};

// Bits of a triangle's flags word (the w of its second vertex row).
constexpr uint32_t kTriHasNormals = 1u;

__device__ __forceinline__ void tri_fetch(const SceneDev& S, uint32_t t, V3* v0, V3* v1, V3* v2, uint32_t* mat,
                                          uint32_t* flags = nullptr) {
    const float4 a = S.tri_v[3 * (size_t)t], b = S.tri_v[3 * (size_t)t + 1], c = S.tri_v[3 * (size_t)t + 2];
    *v0 = v3(a.x, a.y, a.z);
    *v1 = v3(b.x, b.y, b.z);
    *v2 = v3(c.x, c.y, c.z);
    if (mat) *mat = __float_as_uint(a.w);
    if (flags) *flags = __float_as_uint(b.w);
}

// Barycentric weights u (of v0) and v (of v1; v2 takes 1 - u - v) of a
// point p of the triangle's plane, from the normal equations of
// p - v2 = u (v0 - v2) + v (v1 - v2): the Woop transform's u, v convention.
RB_HD void bary_uv(V3 p, V3 v0, V3 v1, V3 v2, float* u, float* v) {
    const V3 e0 = vsub(v0, v2), e1 = vsub(v1, v2), q = vsub(p, v2);
    const float d00 = vdot(e0, e0), d01 = vdot(e0, e1), d11 = vdot(e1, e1);
    const float d20 = vdot(q, e0), d21 = vdot(q, e1);
    const float den = d00 * d11 - d01 * d01;
    *u = den != 0.0f ? (d11 * d20 - d01 * d21) / den : 0.0f;
    *v = den != 0.0f ? (d00 * d21 - d01 * d20) / den : 0.0f;
}

// Shading normal of triangle t at p (the IntersectionBuffer m_shading_normal,
// kepler_dynamic_fetch.h:327-334): vertex normals weighted by p's
// barycentrics, u n0 + v n1 + (1 - u - v) n2, normalised (the reference
// weights n0, n0, n1 and does not normalise).  *ns = ng and false for a
// triangle without normals or where the weighted sum vanishes.  Not
// oriented: it points where the mesh's normals do.
__device__ __forceinline__ bool interp_normal(const SceneDev& S, uint32_t t, uint32_t flags, V3 p, V3 v0, V3 v1, V3 v2,
                                              V3 ng, V3* ns) {
    *ns = ng;
    if (!(flags & kTriHasNormals)) return false;
    const float4 a = S.tri_n[3 * (size_t)t], b = S.tri_n[3 * (size_t)t + 1], c = S.tri_n[3 * (size_t)t + 2];
    float u, v;
    bary_uv(p, v0, v1, v2, &u, &v);
    const float w = 1.0f - u - v;
    const V3 s = v3(a.x * u + b.x * v + c.x * w, a.y * u + b.y * v + c.y * w, a.z * u + b.z * v + c.z * w);
    const float l2 = vdot(s, s);
    if (!(l2 > 0.0f)) return false;
    *ns = vmul(s, 1.0f / sqrtf(l2));
    return true;
}

// The shading normal the passes shade with: interp_normal turned to the
// geometric normal's side.
__device__ __forceinline__ bool shade_normal(const SceneDev& S, uint32_t t, uint32_t flags, V3 p, V3 v0, V3 v1, V3 v2,
                                             V3 ng, V3* ns) {
    const bool sn = interp_normal(S, t, flags, p, v0, v1, v2, ng, ns);
    if (sn && vdot(*ns, ng) < 0.0f) *ns = vneg(*ns);
    return sn;
}

// Importance transport's shading-normal factor at a light sub-path vertex
// (Veach 1997, sec. 5.3): |wo.ns| |wi.ng| / (|wo.ng| |wi.ns|), wo towards the
// light, wi the continuation; 0 where the denominator is.
RB_HD float adjoint_factor(V3 wo, V3 wi, V3 ng, V3 ns) {
    const float num = fabsf(vdot(wo, ns)) * fabsf(vdot(wi, ng));
    const float den = fabsf(vdot(wo, ng)) * fabsf(vdot(wi, ns));
    return den != 0.0f ? num / den : 0.0f;
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

// One sampled direction from a single-layer material: a non-specular BSDF
// gives f and pdf at wi (bsdf_sample); glass gives its throughput factor f
// |cos| / pdf in *f with *weight set; a mirror gives color / |cos| with pdf
// 1 (material_sample_f's lobes, material.cu:98-135).
RB_HD void layer_sample(const rb_material& m, V3 wo, V3 ng, V3 ns, V3 nf, float u1, float u2, bool radiance, V3* wi,
                        V3* f, float* pdf, bool* spec, bool* weight) {
    *weight = false;
    if (!is_delta(m)) {
        bsdf_sample(m, wo, ng, nf, u1, u2, wi, f, pdf);
        *spec = false;
        return;
    }
    *spec = true;
    *pdf = 1.0f;
    if (m.kind == RB_MAT_GLASS) {
        *f = glass_sample(m, wo, ns, u1, radiance, wi);
        *weight = true;
        return;
    }
    // sample_mirror (src/core/reflection.cu:76-92)
    const float c = 2.0f * vdot(nf, wo);
    *wi = vsub(vmul(nf, c), wo);
    const float ac = fabsf(vdot(*wi, nf));
    *f = v3(m.color[0] / ac, m.color[1] / ac, m.color[2] / ac);
}

// ------------------------------------------------------ layered materials --
// RB_MAT_LAYERED (Material's num_valid_layers BSDF layers,
// src/core/material.h:44-52): the equal mixture of its n single-layer
// materials, f = (1/n) sum f_i.  A sampled direction comes from one layer
// picked uniformly (material.cu:80-95); a specular layer's sample keeps its
// own throughput factor ((1/n) delta / (1/n)), a non-specular one has the
// mixture's f and pdf (1/n) sum pdf_i over the non-specular layers at that
// direction.  Light samples and connections evaluate the non-specular
// layers; the light-sample MIS weight uses the pdf of the non-specular
// sampling technique, (1/n_ns) sum pdf_i.  Single-layer materials take the
// plain BSDF code paths unchanged.

RB_HD uint32_t mat_layers(const rb_material& m) { return m.kind == RB_MAT_LAYERED ? (uint32_t)m.param[0] : 1u; }

__device__ __forceinline__ rb_material mat_layer(const SceneDev& S, const rb_material& m, uint32_t k) {
    // a select, not m.param[1 + k]: indexing the register copy at run time
    // would move the whole material to scratch memory
    const float id = k == 0 ? m.param[1] : k == 1 ? m.param[2] : k == 2 ? m.param[3] : m.param[4];
    return S.mats[(uint32_t)id];
}

// every layer specular: no light sample and no connection at this vertex
__device__ __forceinline__ bool mat_delta(const SceneDev& S, const rb_material& m) {
    if (m.kind != RB_MAT_LAYERED) return is_delta(m);
    const uint32_t n = mat_layers(m);
    for (uint32_t k = 0; k < n; k++)
        if (!is_delta(mat_layer(S, m, k))) return false;
    return true;
}

// the material's (non-specular) f for wo -> wi
__device__ __forceinline__ V3 mat_f(const SceneDev& S, const rb_material& m, V3 wo, V3 wi, V3 ng, V3 nf) {
    if (m.kind != RB_MAT_LAYERED) return bsdf_f(m, wo, wi, ng, nf);
    const uint32_t n = mat_layers(m);
    V3 s = v3(0.0f, 0.0f, 0.0f);
    for (uint32_t k = 0; k < n; k++) s = vadd(s, bsdf_f(mat_layer(S, m, k), wo, wi, ng, nf));
    const float fn = (float)n;
    return v3(s.x / fn, s.y / fn, s.z / fn);
}

// sum of the non-specular layers' pdfs at wi, and how many there are
__device__ __forceinline__ float mat_pdf_sum(const SceneDev& S, const rb_material& m, V3 wo, V3 wi, V3 nf,
                                             uint32_t* count) {
    const uint32_t n = mat_layers(m);
    float s = 0.0f;
    uint32_t c = 0;
    for (uint32_t k = 0; k < n; k++) {
        const rb_material L = mat_layer(S, m, k);
        if (is_delta(L)) continue;
        s += bsdf_pdf(L, wo, wi, nf);
        c++;
    }
    *count = c;
    return s;
}

// pdf of the non-specular sampling technique at wi (the light-sample MIS weight)
__device__ __forceinline__ float mat_pdf_ns(const SceneDev& S, const rb_material& m, V3 wo, V3 wi, V3 nf) {
    if (m.kind != RB_MAT_LAYERED) return bsdf_pdf(m, wo, wi, nf);
    uint32_t c;
    const float s = mat_pdf_sum(S, m, wo, wi, nf, &c);
    return c ? s / (float)c : 0.0f;
}

// A direction from the non-specular layers only (one picked uniformly with
// upick), with the material's f and the technique's pdf there.
__device__ __forceinline__ void mat_sample_ns(const SceneDev& S, const rb_material& m, V3 wo, V3 ng, V3 nf, float upick,
                                              float u1, float u2, V3* wi, V3* f, float* pdf) {
    if (m.kind != RB_MAT_LAYERED) {
        bsdf_sample(m, wo, ng, nf, u1, u2, wi, f, pdf);
        return;
    }
    const uint32_t n = mat_layers(m);
    uint32_t c = 0;
    for (uint32_t k = 0; k < n; k++) c += is_delta(mat_layer(S, m, k)) ? 0u : 1u;
    *f = v3(0.0f, 0.0f, 0.0f);
    *pdf = 0.0f;
    if (!c) return;
    uint32_t j = (uint32_t)(upick * (float)c);
    if (j > c - 1) j = c - 1;
    rb_material L = m;
    for (uint32_t k = 0, i = 0; k < n; k++) {
        const rb_material Lk = mat_layer(S, m, k);
        if (is_delta(Lk)) continue;
        if (i++ == j) {
            L = Lk;
            break;
        }
    }
    V3 fl;
    float pl;
    bsdf_sample(L, wo, ng, nf, u1, u2, wi, &fl, &pl);
    if (!(pl > 0.0f)) return;  // the other layers' f may be non-zero where this one's is not
    *f = mat_f(S, m, wo, *wi, ng, nf);
    *pdf = mat_pdf_ns(S, m, wo, *wi, nf);
}

// The next direction from all layers (one picked uniformly with upick):
// layer_sample of that layer, then for a non-specular layer the mixture's f
// and its sampling pdf (1/n) sum pdf_i at wi.
__device__ __forceinline__ void mat_sample(const SceneDev& S, const rb_material& m, V3 wo, V3 ng, V3 ns, V3 nf,
                                           float upick, float u1, float u2, bool radiance, V3* wi, V3* f, float* pdf,
                                           bool* spec, bool* weight) {
    if (m.kind != RB_MAT_LAYERED) {
        layer_sample(m, wo, ng, ns, nf, u1, u2, radiance, wi, f, pdf, spec, weight);
        return;
    }
    const uint32_t n = mat_layers(m);
    uint32_t k = (uint32_t)(upick * (float)n);
    if (k > n - 1) k = n - 1;
    layer_sample(mat_layer(S, m, k), wo, ng, ns, nf, u1, u2, radiance, wi, f, pdf, spec, weight);
    if (*spec || !(*pdf > 0.0f)) return;
    *f = mat_f(S, m, wo, *wi, ng, nf);
    uint32_t c;
    *pdf = mat_pdf_sum(S, m, wo, *wi, nf, &c) / (float)n;
}

}  // namespace rb
