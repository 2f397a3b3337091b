// rb_sample.h -- sampling math for the device passes (and host setup).
//
// Only IEEE-exact operations are used (+ - * /, sqrt, explicit fmaf), so a
// sample computes to the same bits on the GPU and on any IEEE host: no
// library sin/cos (their last-bit results differ between math libraries).
// Built with -ffp-contract=off; every fused multiply-add is written out.
#pragma once
#include <math.h>
#include <string.h>
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

RB_HD uint32_t fbits_(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
RB_HD float bitsf_(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// log2 of a positive finite x: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(t) = 2 (t + t^3/3 + ... + t^9/9), t = (m - 1) / (m + 1),
// |t| <= 0.1716 (truncation < 1e-9); subnormals rescaled by 2^24 first.
RB_HD float log2_pos(float x) {
    uint32_t u = fbits_(x);
    int e = 0;
    if (u < 0x00800000u) {
        x = x * 16777216.0f;
        u = fbits_(x);
        e = -24;
    }
    e += (int)(u >> 23) - 127;
    float m = bitsf_((u & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float t = (m - 1.0f) / (m + 1.0f);
    const float t2 = t * t;
    float q = fmaf(t2, 0.222222222f, 0.285714286f);
    q = fmaf(t2, q, 0.4f);
    q = fmaf(t2, q, 0.666666667f);
    q = fmaf(t2, q, 2.0f);
    return fmaf(t * q, 1.44269504f, (float)e);
}

// 2^y: k = nearest integer, f = y - k in [-1/2, 1/2], 2^f = e^(f ln 2) from
// its Taylor series to degree 7 (truncation < 6e-9), scaled by 2^k through
// the exponent field.
RB_HD float exp2_series(float y) {
    if (y != y) return y;
    if (y >= 128.0f) return bitsf_(0x7F800000u);
    if (y < -150.0f) return 0.0f;
    const float k = floorf(y + 0.5f);
    const float z = (y - k) * 0.693147181f;
    float q = fmaf(z, 1.98412698e-4f, 1.38888889e-3f);
    q = fmaf(z, q, 8.33333333e-3f);
    q = fmaf(z, q, 4.16666667e-2f);
    q = fmaf(z, q, 0.166666667f);
    q = fmaf(z, q, 0.5f);
    q = fmaf(z, q, 1.0f);
    q = fmaf(z, q, 1.0f);
    int ki = (int)k;
    if (ki > 127) {
        q = q * 2.0f;
        ki = 127;
    }
    if (ki < -126) {
        q = q * 5.96046448e-8f;  // 2^-24, exact
        ki += 24;
    }
    return q * bitsf_((uint32_t)(ki + 127) << 23);
}

// x^y for x >= 0 as 2^(y log2 x): the same bits on every IEEE device (no
// math-library call); relative error ~1e-6 for |y log2 x| up to ~30.
RB_HD float pow_series(float x, float y) {
    if (x > 0.0f) return exp2_series(y * log2_pos(x));
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : bitsf_(0x7F800000u));
    return bitsf_(0x7FC00000u);
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
