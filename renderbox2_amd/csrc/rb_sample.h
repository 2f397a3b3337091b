// rb_sample.h -- sampling math for the device passes (and host setup).
//
// Only IEEE-exact operations are used (+ - * /, sqrt, explicit fmaf), so a
// sample computes to the same bits on the GPU and on any IEEE host: no
// library sin/cos (their last-bit results differ between math libraries).
// Built with -ffp-contract=off; every fused multiply-add is written out.
#pragma once
#include <math.h>
#include <stdint.h>

#include "rb_rng.h"

namespace rb {

struct V3 {
    float x, y, z;
};
RB_HD V3 v3(float x, float y, float z) {
    V3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
RB_HD V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RB_HD V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
RB_HD V3 vmul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
RB_HD V3 vmulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
RB_HD V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
RB_HD float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RB_HD V3 vcross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
// helper_math normalize with an exactly rounded reciprocal square root.
RB_HD V3 vnorm(V3 a) { return vmul(a, 1.0f / sqrtf(vdot(a, a))); }

// sin and cos of 2*pi*u for u in [0, 1): quadrant from the nearest quarter
// turn (exact: 4u and 4u - k are exact in binary32), then Taylor series of
// degree 9 (sin) and 10 (cos) on |x| <= pi/4, truncation error < 2e-9.
RB_HD void sincos_2pi(float u, float* s, float* c) {
    const float a = u * 4.0f;
    const int k = (int)(a + 0.5f);
    const float r = a - (float)k;
    const float x = r * 1.57079632679489662f;
    const float x2 = x * x;
    float ps = fmaf(x2, 2.75573192e-06f, -1.98412698e-04f);
    ps = fmaf(x2, ps, 8.33333333e-03f);
    ps = fmaf(x2, ps, -1.66666667e-01f);
    const float sn = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, -2.75573192e-07f, 2.48015873e-05f);
    pc = fmaf(x2, pc, -1.38888889e-03f);
    pc = fmaf(x2, pc, 4.16666667e-02f);
    pc = fmaf(x2, pc, -0.5f);
    const float cs = fmaf(x2, pc, 1.0f);
    switch (k & 3) {
        case 0: *s = sn; *c = cs; break;
        case 1: *s = cs; *c = -sn; break;
        case 2: *s = -sn; *c = -cs; break;
        default: *s = -cs; *c = sn; break;
    }
}

// Orthonormal basis around unit n (Duff, Burgess, Christensen, Hery, Kensler,
// Liani, Villemin, "Building an Orthonormal Basis, Revisited", JCGT 2017).
RB_HD void onb(V3 n, V3* b1, V3* b2) {
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    *b1 = v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    *b2 = v3(b, sign + n.y * n.y * a, -n.y);
}

// Cosine-weighted direction about unit n from (u1, u2) in [0,1)^2
// (Malley's method: uniform disk point lifted to the hemisphere).
RB_HD V3 cosine_hemisphere(V3 n, float u1, float u2) {
    float s, c;
    sincos_2pi(u2, &s, &c);
    const float r = sqrtf(u1);
    const float x = r * c, y = r * s;
    const float zz = 1.0f - u1;
    const float z = sqrtf(zz > 0.0f ? zz : 0.0f);
    V3 b1, b2;
    onb(n, &b1, &b2);
    return vadd(vadd(vmul(b1, x), vmul(b2, y)), vmul(n, z));
}

// Uniform point on a triangle (uniform_sample_triangle, src/core/montecarlo.cu:157-164).
RB_HD V3 sample_triangle(V3 v0, V3 v1, V3 v2, float u1, float u2) {
    const float su1 = sqrtf(u1);
    const float u = 1.0f - su1;
    const float v = u2 * su1;
    return vadd(vadd(vmul(v0, u), vmul(v1, v)), vmul(v2, 1.0f - u - v));
}

// compute_normal (src/core/primitives.h:241-245).
RB_HD V3 tri_normal(V3 v0, V3 v1, V3 v2) { return vnorm(vcross(vsub(v1, v0), vsub(v2, v0))); }

}  // namespace rb
