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

}  // namespace rb
