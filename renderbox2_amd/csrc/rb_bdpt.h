// rb_bdpt.h -- bidirectional path tracing with a light-vertex cache (LVC).
//
// Restates the reference's BDPT-LVC (src/integrators/integrator_bdpt_lvc_sk.cu:47-435,
// wavefront form integrator_bdpt_lvc_mk.cu) as wavefront passes:
//  light pass  : light sub-paths from uniformly chosen area lights (point
//                uniform on the triangle, direction uniform over the
//                emitting hemisphere); every non-emitter hit is stored in the
//                cache at slot (path, depth) -- fixed slots, so the cache is
//                identical run to run (the reference appends with atomics);
//  camera pass : per camera vertex, direct light (NEE + BSDF sample with
//                MIS, as the path tracer), `connections` shadow-ray
//                connections to cache vertices, then the next bounce; any-hit
//                and closest-hit rays of a bounce go through one persistent
//                launch each, live paths are compacted.
// Weights: every path is weighted by 1 / (number of techniques that can
// build it): the direct-light pair counts as one technique and each
// connection whose two end vertices are non-specular, and whose camera and
// light sub-paths fit max_depth and the cache depth, as one more; with no
// connections the weights are all 1 and the camera pass is the path tracer.
// The reference weights a connection by 1 / ((bounce + 1) + lvc_depth)
// whatever the path (integrator_bdpt_lvc_sk.cu:379-384) and adds the direct
// term at full weight, which counts paths more than once.
// Connection estimator: pick a light path j uniformly, then one of its n_j
// stored vertices uniformly, and scale by n_j / connections (unbiased).
#pragma once
#include "rb_shade.h"
#include "../../include/rb_trace_device.h"

namespace rb {

// Light-vertex cache: one 64-byte record per vertex (one cache line per
// random gather by a connection), rows
//   [0] p.xyz | 1 when row 4 holds an interpolated shading normal, else 0
//   [1] geometric normal | material id bits
//   [2] direction towards the previous light vertex | specular mask bits
//   [3] light sub-path throughput arriving at the vertex
//   [4] (scenes with shading normals only: 80-byte records) shading normal
struct LvcBufs {
    float4* v;        // `rows` rows per vertex, vertex `k` of light path `j` at slot j * depth + k
    uint32_t* count;  // vertices stored per light path
    uint32_t depth;   // slots per light path
    uint32_t paths;
    uint32_t rows;    // 4, or 5 in scenes with shading normals
};

struct LightPath {
    float4* ray;    // 2 per path
    float4* alpha;
    uint32_t* mask;
    uint32_t* q;    // the emitted paths, in queue order
    uint32_t* cnt;  // [0]: their number
    int2* hit;      // depth 0's hit of queue entry q
};

// Queue kernels (shade, connect, path and light-path init / extension) run
// kQueueBlock-thread workgroups whose loops take the same number of trips in
// every thread of a workgroup, as block_reserve needs.
constexpr uint32_t kQueueBlock = 512;

// Reserves c (0..31) consecutive slots of the queue whose length is
// *counter for the calling thread and returns the first (meaningless when
// c == 0).  The whole workgroup reserves at once: wave prefix sums from
// bit-plane ballots, wave offsets through LDS, one global atomic per call.
// Appends to one queue counter from every wave of the chip serialise where
// the device-scope atomics are performed (~10 ns each, one per wave made the
// queue kernels atomic-bound); per workgroup they are kQueueBlock / 64 times
// fewer.  Every thread of the workgroup must call it.
__device__ __forceinline__ uint32_t block_reserve(uint32_t* counter, uint32_t c) {
    __shared__ uint32_t s_off[kQueueBlock / 64 + 1];
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t pre = 0, tot = 0;
    for (uint32_t b = 0; b < 5; b++) {
        const uint64_t m = __ballot((c >> b) & 1u);
        pre += (uint32_t)__popcll(m & lt) << b;
        tot += (uint32_t)__popcll(m) << b;
    }
    if (lane == 0) s_off[wave] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
        for (uint32_t w = 0; w < nw; w++) {
            const uint32_t t = s_off[w];
            s_off[w] = sum;
            sum += t;
        }
        s_off[kQueueBlock / 64] = sum ? atomicAdd(counter, sum) : 0u;
    }
    __syncthreads();
    const uint32_t base = s_off[kQueueBlock / 64] + s_off[wave] + pre;
    __syncthreads();  // s_off is reused by the next call
    return base;
}

// End of a workgroup-uniform loop over n items.
__device__ __forceinline__ uint32_t uniform_end(uint32_t n) { return (n + blockDim.x - 1) / blockDim.x * blockDim.x; }

// uniform_sample_hemisphere (src/core/montecarlo.cu:116-124) about unit n.
__device__ __forceinline__ V3 uniform_hemisphere(V3 n, float u1, float u2) {
    float s, c;
    sincos_2pi(u2, &s, &c);
    const float z = u1;
    const float rr = 1.0f - z * z;
    const float r = sqrtf(rr > 0.0f ? rr : 0.0f);
    V3 b1, b2;
    onb(n, &b1, &b2);
    return vadd(vadd(vmul(b1, r * c), vmul(b2, r * s)), vmul(n, z));
}

// 1 / (number of techniques that can build a path z_1 .. z_n; z_0 is the
// camera, z_n the point on the light, mask bit i-1 = z_i specular, z_n
// never).  Techniques, each counted only when its sub-paths fit the depth
// limits (camera vertices exist for bounces 0..max_depth, the cache holds
// light vertices at depths 0..lvc_depth-1):
//  * direct light: light sample + MIS BSDF sample at a diffuse z_(n-1)
//    (bounce n-2), or the emitter hit at bounce n-1 when z_(n-1) is
//    specular or n = 1;
//  * connection edge (z_i, z_(i+1)), 1 <= i <= n-2, both ends diffuse,
//    camera end at bounce i-1 <= max_depth, light end at depth n-i-2 <
//    lvc_depth.
// n <= 32 (the renderer bounds max_depth + lvc_depth).
RB_HD float technique_weight(uint32_t mask, uint32_t n, uint32_t max_depth, uint32_t lvc_depth) {
    uint32_t c = 1;
    if (n >= 2) c = ((mask >> (n - 2)) & 1u) ? (n - 1 <= max_depth) : (n - 2 <= max_depth);
    if (n >= 3) {
        uint32_t ok = ~mask & ~(mask >> 1) & ((1u << (n - 2)) - 1u);  // bit i-1: edge i
        const uint32_t lo = n - 2 > lvc_depth ? n - 2 - lvc_depth : 0u;
        ok &= ~((1u << lo) - 1u);
        if (max_depth < 31) ok &= (2u << max_depth) - 1u;
        c += (uint32_t)__builtin_popcount(ok);
    }
    return c ? 1.0f / (float)c : 0.0f;
}

__global__ __launch_bounds__(kQueueBlock) void k_light_emit(uint32_t npaths, SceneDev S, RngArg rng_a, LightPath P,
                                                             LvcBufs C) {
    const Rng rng = rng_a.get();
    const uint32_t end = uniform_end(npaths);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < end; j += gridDim.x * blockDim.x) {
        const bool valid = j < npaths;
        const bool live = valid && S.n_lights > 0;  // no emitters: no light paths, empty cache
        if (valid) C.count[j] = 0;
        if (live) {
            const U4 r = rng.draw(j, 0, 0, 0);
            const U4 r2 = rng.draw(j, 0, 1, 0);
            uint32_t lid = (uint32_t)(u01(r.x) * (float)S.n_lights);
            if (lid > S.n_lights - 1) lid = S.n_lights - 1;
            const uint32_t lt = S.light_tris[lid];
            V3 l0, l1, l2;
            uint32_t lmid;
            tri_fetch(S, lt, &l0, &l1, &l2, &lmid);
            const rb_material lm = S.mats[lmid];
            const V3 ln = tri_normal(l0, l1, l2);
            const float area = tri_area(l0, l1, l2);
            const V3 pl = sample_triangle(l0, l1, l2, u01(r.y), u01(r.z));
            const V3 dir = uniform_hemisphere(ln, u01(r.w), u01(r2.x));
            const float cth = fabsf(vdot(ln, dir));
            // DiffuseAreaLight::Le (src/core/lights.cu:68-79): pdf over area x
            // projected solid angle; times the light count (uniform choice)
            const float pdf = (1.0f / area) * 0.159154943091895336f / cth;
            const float k = (float)S.n_lights / pdf;
            P.alpha[j] = make_float4(lm.color[0] * k, lm.color[1] * k, lm.color[2] * k, 0.f);
            P.ray[2 * (size_t)j] = make_float4(pl.x, pl.y, pl.z, 1.0e-4f);
            P.ray[2 * (size_t)j + 1] = make_float4(dir.x, dir.y, dir.z, 3.402823466e+38f);
            P.mask[j] = 0;
        }
        const uint32_t q = block_reserve(&P.cnt[0], live ? 1u : 0u);
        if (live) P.q[q] = j;
    }
}

// One light vertex (kernel_sk_light_trace's body, integrator_bdpt_lvc_sk.cu:
// 47-168): light path j's ray (o, dd) hit h at depth d.  A non-emitter hit
// is stored at slot (j, d) with the throughput a arriving there; then, below
// the cache depth, a BSDF sample with roulette (continue probability
// min(1, luminance)) gives the next ray, throughput and specular mask in
// place and returns true.  Emitters end the path.
__device__ __forceinline__ bool light_vertex(const SceneDev& S, const Rng& rng, const LvcBufs& C, uint32_t j, uint32_t d,
                                             int2 h, float4* o, float4* dd, float4* a, uint32_t* mask) {
    const float t = __int_as_float(h.y);
    const V3 dv = v3(dd->x, dd->y, dd->z);
    const V3 p = vadd(v3(o->x, o->y, o->z), vmul(dv, t));
    V3 v0, v1, v2;
    uint32_t mid, tf;
    tri_fetch(S, (uint32_t)h.x, &v0, &v1, &v2, &mid, &tf);
    const rb_material m = S.mats[mid];
    if (m.kind == RB_MAT_EMITTER) return false;
    const V3 ng = tri_normal(v0, v1, v2);
    V3 ns;
    const bool sn = shade_normal(S, (uint32_t)h.x, tf, p, v0, v1, v2, ng, &ns);
    const V3 wo = vneg(dv);
    // the stored vertex is specular when every layer is; the path continues
    // with the sampled lobe's kind
    const uint32_t dbit = mat_delta(S, m) ? 1u << d : 0u;
    const size_t slot = (size_t)j * C.depth + d;
    float4* rec = C.v + C.rows * slot;
    rec[0] = make_float4(p.x, p.y, p.z, __uint_as_float(sn ? 1u : 0u));
    rec[1] = make_float4(ng.x, ng.y, ng.z, __uint_as_float(mid));
    rec[2] = make_float4(wo.x, wo.y, wo.z, __uint_as_float(*mask | dbit));
    rec[3] = *a;
    if (C.rows > 4) rec[4] = make_float4(ns.x, ns.y, ns.z, 0.f);
    C.count[j] = d + 1;
    if (d + 1 >= C.depth) return false;
    const U4 r = rng.draw(j, d + 1, 0, 0);
    const V3 nf = vdot(ng, wo) >= 0.0f ? ns : vneg(ns);
    V3 wi, f;
    float pdf;
    bool spec, weight;  // weight: f already holds f |cos| / pdf
    mat_sample(S, m, wo, ng, ns, nf, u01(r.w), u01(r.x), u01(r.y), false, &wi, &f, &pdf, &spec, &weight);
    if (is_black(f) || !(pdf > 0.0f)) return false;
    const float s = weight ? 1.0f : fabsf(vdot(ns, wi));
    V3 tmp = v3(f.x * s / pdf, f.y * s / pdf, f.z * s / pdf);
    // importance transport: shading-normal adjoint factor
    if (sn) tmp = vmul(tmp, adjoint_factor(wo, wi, ng, ns));
    const float lum = luminance(tmp);
    const float cp = lum < 1.0f ? lum : 1.0f;
    if (u01(r.z) > cp) return false;
    const float4 na = make_float4(a->x * (tmp.x / cp), a->y * (tmp.y / cp), a->z * (tmp.z / cp), 0.f);
    if (na.x == 0.f && na.y == 0.f && na.z == 0.f) return false;
    *a = na;
    if (spec) *mask |= 1u << d;
    *o = make_float4(p.x, p.y, p.z, 1.0e-3f * t);
    *dd = make_float4(wi.x, wi.y, wi.z, 3.402823466e+38f);
    return true;
}

// The light paths after their first trace, one thread per path: depth 0's
// hit comes from the launch that traced the light rays with the camera rays
// (queue 0, hit q of entry q); every later depth's ray is traced by the
// thread itself with rb_trace_one, whose (id, t) equal the persistent
// kernel's bit for bit.  One launch for all depths replaces a trace launch
// per depth, whose ramp-up and tail dominated the thinning later depths.
// Optional counters: rays traced (ray_counter) and rays per depth
// (depth_counter[d], d >= 1).
__global__ __launch_bounds__(256) void k_light_walk(SceneDev S, RngArg rng_a, LightPath P, LvcBufs C, rb_bvh_dev bv,
                                                    unsigned long long* ray_counter,
                                                    unsigned long long* depth_counter) {
    const Rng rng = rng_a.get();
    const uint32_t count = P.cnt[0];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x) {
        const uint32_t j = P.q[q];
        int2 h = P.hit[q];
        float4 o = P.ray[2 * (size_t)j], dd = P.ray[2 * (size_t)j + 1];
        float4 a = P.alpha[j];
        uint32_t mask = P.mask[j];
        for (uint32_t d = 0; h.x >= 0 && light_vertex(S, rng, C, j, d, h, &o, &dd, &a, &mask);) {
            d++;
            const rb_ray r = {o.x, o.y, o.z, o.w, dd.x, dd.y, dd.z, dd.w};
            const rb_hit hh = rb_trace_one(bv, r, RB_CLOSEST);
            h = make_int2(hh.id, __float_as_int(hh.t));
            if (ray_counter) atomicAdd(ray_counter, 1ull);
            if (depth_counter) atomicAdd(depth_counter + d, 1ull);
        }
    }
}

// Gather the rays of the emitted light paths into a dense buffer for tracing.
__global__ void k_light_gather(LightPath P, float4* rays) {
    const uint32_t count = P.cnt[0];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x) {
        const uint32_t j = P.q[q];
        rays[2 * (size_t)q] = P.ray[2 * (size_t)j];
        rays[2 * (size_t)q + 1] = P.ray[2 * (size_t)j + 1];
    }
}

}  // namespace rb
