/*
 * rb_trace_device.h -- header-only device-side single-ray traversal for
 * HIP kernels that trace from inside their own code.
 *
 * Replaces the reference's device entry point
 *   extern "C" __device__ bool trace_ray(const Ray&, bool anyhit, Intersection*,
 *       cudaTextureObject_t nodes, tris, triIndices, SceneBuffer)
 * (src/accelerators/sbvh/cudatracerkernels.h:132-141, body
 * src/accelerators/sbvh/cuda_tracer_functions.cu:42-279) with
 *   __device__ rb_hit rb_trace_one(const rb_bvh_dev&, const rb_ray&, uint32_t flags)
 * over the MI355X node layout of a resident BVH (rb_bvh_device_view()).
 *
 * The ray walks the tree in the same order as the persistent kernel behind
 * rb_trace (default, non-speculative leaf order) with the same fp32
 * arithmetic, so (id, t) equal rb_trace's bit for bit; ties at equal t go to
 * the lowest original triangle id.  The stack lives in private memory (96
 * entries), so this is for callers tracing a few rays per thread; batches
 * belong to rb_trace / rb_trace_batch.
 *
 * Compile with -ffp-contract=off for results identical to rb_trace.
 */
#ifndef RB_TRACE_DEVICE_H
#define RB_TRACE_DEVICE_H

#include <hip/hip_runtime.h>

#include "rb.h"

#define RB_DEV_STACK 96
#define RB_DEV_SENTINEL 0x76543210

/* Node record: four 16-byte rows -- child 0 x/y slabs, child 1 x/y slabs,
 * both z slabs, links (child 0, child 1, counts).  A link >= 0 is an inner
 * node index; a leaf is ~(first << 5 | count) with `count` Woop references
 * of 3 rows each starting at reference `first`. */
__device__ __forceinline__ rb_hit rb_trace_one(const rb_bvh_dev& b, const rb_ray& ray, uint32_t flags) {
    const bool any = (flags & RB_ANY) != 0;
    const float ox = ray.ox, oy = ray.oy, oz = ray.oz, tmin = ray.tmin;
    const float dx = ray.dx, dy = ray.dy, dz = ray.dz;
    float hitT = ray.tmax;
    const float ooeps = 0x1p-80f;
    const float idx = 1.0f / (fabsf(dx) > ooeps ? dx : copysignf(ooeps, dx));
    const float idy = 1.0f / (fabsf(dy) > ooeps ? dy : copysignf(ooeps, dy));
    const float idz = 1.0f / (fabsf(dz) > ooeps ? dz : copysignf(ooeps, dz));
    const float oodx = ox * idx, oody = oy * idy, oodz = oz * idz;
    const float4* nodes = reinterpret_cast<const float4*>(b.nodes);
    const float4* tris = reinterpret_cast<const float4*>(b.tris);
    int stack[RB_DEV_STACK];
    int sp = 0;
    stack[0] = RB_DEV_SENTINEL;
    int nodeAddr = 0, leafAddr = 0, hitIndex = -1;
    /* the upload refuses trees deep enough to fill the stack; a push past
       it is dropped and flagged (rb_trace_errors) */
#define RB_DEV_PUSH(v)                                               \
    do {                                                             \
        if (sp + 1 < RB_DEV_STACK) {                                 \
            ++sp;                                                    \
            stack[sp] = (v);                                         \
        } else if (b.errors) {                                       \
            atomicOr(b.errors, (uint32_t)RB_TRACE_ERR_STACK);        \
        }                                                            \
    } while (0)
#define RB_DEV_POP(dst)      \
    do {                     \
        (dst) = stack[sp];   \
        --sp;                \
    } while (0)
    while (nodeAddr != RB_DEV_SENTINEL) {
        /* inner nodes until a leaf is reached */
        while ((uint32_t)nodeAddr < (uint32_t)RB_DEV_SENTINEL && leafAddr >= 0) {
            const float4* np = nodes + 4 * (size_t)nodeAddr;
            const float4 n0 = np[0], n1 = np[1], nz = np[2];
            const int4 lk = *reinterpret_cast<const int4*>(np + 3);
            /* keep the four loads of the record together (the link load would
               otherwise be sunk below the slab tests: a second round trip) */
            asm volatile("" ::"v"(n0.x), "v"(n0.y), "v"(n0.z), "v"(n0.w), "v"(n1.x), "v"(n1.y), "v"(n1.z),
                         "v"(n1.w), "v"(nz.x), "v"(nz.y), "v"(nz.z), "v"(nz.w), "v"(lk.x), "v"(lk.y));
            const float c0lox = fmaf(n0.x, idx, -oodx), c0hix = fmaf(n0.y, idx, -oodx);
            const float c0loy = fmaf(n0.z, idy, -oody), c0hiy = fmaf(n0.w, idy, -oody);
            const float c0loz = fmaf(nz.x, idz, -oodz), c0hiz = fmaf(nz.y, idz, -oodz);
            const float c1loz = fmaf(nz.z, idz, -oodz), c1hiz = fmaf(nz.w, idz, -oodz);
            const float c1lox = fmaf(n1.x, idx, -oodx), c1hix = fmaf(n1.y, idx, -oodx);
            const float c1loy = fmaf(n1.z, idy, -oody), c1hiy = fmaf(n1.w, idy, -oody);
            const float c0min = fmaxf(fmaxf(fminf(c0lox, c0hix), fminf(c0loy, c0hiy)), fmaxf(fminf(c0loz, c0hiz), tmin));
            const float c0max = fminf(fminf(fmaxf(c0lox, c0hix), fmaxf(c0loy, c0hiy)), fminf(fmaxf(c0loz, c0hiz), hitT));
            const float c1min = fmaxf(fmaxf(fminf(c1lox, c1hix), fminf(c1loy, c1hiy)), fmaxf(fminf(c1loz, c1hiz), tmin));
            const float c1max = fminf(fminf(fmaxf(c1lox, c1hix), fmaxf(c1loy, c1hiy)), fminf(fmaxf(c1loz, c1hiz), hitT));
            const bool swp = c1min < c0min;
            const bool t0 = c0max >= c0min;
            const bool t1 = c1max >= c1min;
            int far = lk.y;
            if (!t0 && !t1) {
                RB_DEV_POP(nodeAddr);
            } else {
                nodeAddr = t0 ? lk.x : lk.y;
                if (t0 && t1) {
                    if (swp) {
                        const int tmp = nodeAddr;
                        nodeAddr = far;
                        far = tmp;
                    }
                    RB_DEV_PUSH(far);
                }
            }
            if (nodeAddr < 0) {
                leafAddr = nodeAddr;
                RB_DEV_POP(nodeAddr);
            }
        }
        /* the leaf, then any leaves popped right after it */
        while (leafAddr < 0) {
            const uint32_t l = ~(uint32_t)leafAddr;
            const uint32_t first = l >> 5, cnt = l & 31u;
            for (uint32_t k = 0; k < cnt; k++) {
                const float4* w = tris + 3 * (size_t)(first + k);
                const float4 v0 = w[0], v1 = w[1], v2 = w[2];
                const float Oz = fmaf(-oz, v0.z, fmaf(-oy, v0.y, fmaf(-ox, v0.x, v0.w)));
                const float invDz = 1.0f / fmaf(dz, v0.z, fmaf(dy, v0.y, dx * v0.x));
                const float t = Oz * invDz;
                if (t > tmin && (t < hitT || (hitIndex != -1 && t == hitT))) {
                    const float Ox = fmaf(oz, v1.z, fmaf(oy, v1.y, fmaf(ox, v1.x, v1.w)));
                    const float Dx = fmaf(dz, v1.z, fmaf(dy, v1.y, dx * v1.x));
                    const float u = fmaf(t, Dx, Ox);
                    if (u >= 0.0f) {
                        const float Oy = fmaf(oz, v2.z, fmaf(oy, v2.y, fmaf(ox, v2.x, v2.w)));
                        const float Dy = fmaf(dz, v2.z, fmaf(dy, v2.y, dx * v2.x));
                        const float v = fmaf(t, Dy, Oy);
                        if (v >= 0.0f && u + v <= 1.0f) {
                            const bool take = t < hitT || b.index[first + k] < b.index[hitIndex];
                            if (take) {
                                hitT = t;
                                hitIndex = (int)(first + k);
                                if (any) {
                                    nodeAddr = RB_DEV_SENTINEL;
                                    break;
                                }
                            }
                        }
                    }
                }
            }
            leafAddr = nodeAddr;
            if (nodeAddr < 0) RB_DEV_POP(nodeAddr);
        }
    }
#undef RB_DEV_PUSH
#undef RB_DEV_POP
    rb_hit h;
    h.id = hitIndex < 0 ? -1 : (int32_t)b.index[hitIndex];
    h.t = hitT;
    return h;
}

#endif /* RB_TRACE_DEVICE_H */
