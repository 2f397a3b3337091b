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
 *   __device__ bool rb_trace_one_isect(const rb_bvh_dev&, const rb_ray&, uint32_t flags, rb_isect_one*)
 * over the node records of a resident BVH (rb_bvh_device_view()).
 *
 * A closest-hit ray walks the wide (4-child) records when the view has them,
 * else the binary records in the same order as the persistent kernel behind
 * rb_trace (default, non-speculative leaf order); any hit walks the binary
 * records.  Same fp32 arithmetic, so (id, t) equal rb_trace's bit for bit;
 * ties at equal t go to the lowest original triangle id.  The stack lives in
 * private memory (96 entries), so this is for callers tracing a few rays per
 * thread; batches belong to rb_trace / rb_trace_batch.
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
    /* closest hit over the wide records: the same (id, t), about half the
       record fetches (every leaf the binary walk tests is tested) */
    const float4* wnodes = any ? nullptr : reinterpret_cast<const float4*>(b.wnodes);
    while (nodeAddr != RB_DEV_SENTINEL) {
        /* wide records until a leaf is reached */
        while (wnodes && (uint32_t)nodeAddr < (uint32_t)RB_DEV_SENTINEL && leafAddr >= 0) {
            /* record: origin p.xyz | biased exponents (byte a: e + 127 of axis
               a); the children's 8-bit steps for lo.x, hi.x, lo.y, hi.y, lo.z,
               hi.z (byte c = child c); four links (0x7fffffff: no child) */
            const float4* np = wnodes + 4 * (size_t)nodeAddr;
            const float4 r0 = np[0];
            const uint4 r1 = *reinterpret_cast<const uint4*>(np + 1);
            const uint4 r2 = *reinterpret_cast<const uint4*>(np + 2);
            const uint4 r3 = *reinterpret_cast<const uint4*>(np + 3);
            asm volatile("" ::"v"(r0.x), "v"(r1.x), "v"(r2.x), "v"(r3.x));
            const uint32_t eb = __float_as_uint(r0.w);
            const float sx = __uint_as_float((eb & 0xFFu) << 23);
            const float sy = __uint_as_float(((eb >> 8) & 0xFFu) << 23);
            const float sz = __uint_as_float(((eb >> 16) & 0xFFu) << 23);
            const int lk[4] = {(int)r2.z, (int)r2.w, (int)r3.x, (int)r3.y};
            float key[4];
            int lnk[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                /* a plane is fmaf(step, 2^e, p), a box holding the exact one */
#define RB_DEV_Q(w, s, p) fmaf((float)(((w) >> (8 * c)) & 0xFFu), s, p)
                const float lx = RB_DEV_Q(r1.x, sx, r0.x), hx = RB_DEV_Q(r1.y, sx, r0.x);
                const float ly = RB_DEV_Q(r1.z, sy, r0.y), hy = RB_DEV_Q(r1.w, sy, r0.y);
                const float lz = RB_DEV_Q(r2.x, sz, r0.z), hz = RB_DEV_Q(r2.y, sz, r0.z);
#undef RB_DEV_Q
                const float clox = fmaf(lx, idx, -oodx), chix = fmaf(hx, idx, -oodx);
                const float cloy = fmaf(ly, idy, -oody), chiy = fmaf(hy, idy, -oody);
                const float cloz = fmaf(lz, idz, -oodz), chiz = fmaf(hz, idz, -oodz);
                const float cmin = fmaxf(fmaxf(fminf(clox, chix), fminf(cloy, chiy)), fmaxf(fminf(cloz, chiz), tmin));
                const float cmax = fminf(fminf(fmaxf(clox, chix), fmaxf(cloy, chiy)), fminf(fmaxf(cloz, chiz), hitT));
                const bool h = cmax >= cmin && lk[c] != 0x7fffffff;
                key[c] = h ? fminf(cmin, 0x1.fffffep127f) : __builtin_inff();
                lnk[c] = lk[c];
            }
            /* hit children nearest first, misses last */
#define RB_DEV_CS(i, j)                                        \
    if (key[j] < key[i]) {                                     \
        const float tk = key[i];                               \
        key[i] = key[j];                                       \
        key[j] = tk;                                           \
        const int tl = lnk[i];                                 \
        lnk[i] = lnk[j];                                       \
        lnk[j] = tl;                                           \
    }
            RB_DEV_CS(0, 1) RB_DEV_CS(2, 3) RB_DEV_CS(0, 2) RB_DEV_CS(1, 3) RB_DEV_CS(1, 2)
#undef RB_DEV_CS
            if (key[0] == __builtin_inff()) {
                RB_DEV_POP(nodeAddr);
            } else {
                if (key[3] != __builtin_inff()) RB_DEV_PUSH(lnk[3]);
                if (key[2] != __builtin_inff()) RB_DEV_PUSH(lnk[2]);
                if (key[1] != __builtin_inff()) RB_DEV_PUSH(lnk[1]);
                nodeAddr = lnk[0];
            }
            if (nodeAddr < 0) {
                leafAddr = nodeAddr;
                RB_DEV_POP(nodeAddr);
            }
        }
        /* binary records until a leaf is reached */
        while (!wnodes && (uint32_t)nodeAddr < (uint32_t)RB_DEV_SENTINEL && leafAddr >= 0) {
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

/* One intersection record: the fields of the reference's Intersection
 * (src/core/intersection.h:42-60) that trace_ray writes, plus t and the
 * barycentrics. */
typedef struct rb_isect_one {
    int32_t primitive_id;          /* original triangle id, -1 on a miss */
    uint32_t material_id;
    float t;                       /* hit distance (the ray's tmax on a miss) */
    float position[3];             /* o + t d */
    float incoming_direction[3];   /* -d */
    float geometric_normal[3];     /* normalize(cross(v1 - v0, v2 - v0)) */
    float shading_normal[3];       /* vertex normals at the hit, normalised (Ng without them) */
    float uv[2];                   /* barycentric weights of v0 and v1 */
    float epsilon;                 /* 1e-3 t */
    uint8_t front;                 /* Ng . d < 0 */
    uint8_t intersected;
} rb_isect_one;

/* The record of hit h of ray (the result-write block of trace_ray,
 * cuda_tracer_functions.cu:236-276): intersected and t always; for a hit,
 * the rest from the scene rows of the view (b.tri_v: a BVH uploaded with its
 * scene).  Without them only the position, direction, epsilon and ids are
 * written.  The same arithmetic as rb_trace_fill, so the same bits. */
__device__ __forceinline__ void rb_fill_one(const rb_bvh_dev& b, const rb_ray& ray, rb_hit h, rb_isect_one* o) {
    o->intersected = h.id >= 0 ? 1 : 0;
    o->t = h.t;
    o->primitive_id = h.id;
    if (h.id < 0) return;
    const float t = h.t;
    const float px = ray.ox + ray.dx * t, py = ray.oy + ray.dy * t, pz = ray.oz + ray.dz * t;
    o->position[0] = px;
    o->position[1] = py;
    o->position[2] = pz;
    o->incoming_direction[0] = -ray.dx;
    o->incoming_direction[1] = -ray.dy;
    o->incoming_direction[2] = -ray.dz;
    o->epsilon = 1.0e-3f * t;
    o->material_id = 0;
    if (!b.tri_v) return;
    const float4* tv = reinterpret_cast<const float4*>(b.tri_v) + 3 * (size_t)h.id;
    const float4 a = tv[0], bb = tv[1], c = tv[2];
    o->material_id = __float_as_uint(a.w);
    /* geometric normal */
    const float e1x = bb.x - a.x, e1y = bb.y - a.y, e1z = bb.z - a.z;
    const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
    const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
    const float inv = 1.0f / sqrtf(cx * cx + cy * cy + cz * cz);
    const float ngx = cx * inv, ngy = cy * inv, ngz = cz * inv;
    o->geometric_normal[0] = ngx;
    o->geometric_normal[1] = ngy;
    o->geometric_normal[2] = ngz;
    o->front = ngx * ray.dx + ngy * ray.dy + ngz * ray.dz < 0.0f ? 1 : 0;
    /* barycentrics of p: p - v2 = u (v0 - v2) + v (v1 - v2) */
    const float f0x = a.x - c.x, f0y = a.y - c.y, f0z = a.z - c.z;
    const float f1x = bb.x - c.x, f1y = bb.y - c.y, f1z = bb.z - c.z;
    const float qx = px - c.x, qy = py - c.y, qz = pz - c.z;
    const float d00 = f0x * f0x + f0y * f0y + f0z * f0z, d01 = f0x * f1x + f0y * f1y + f0z * f1z;
    const float d11 = f1x * f1x + f1y * f1y + f1z * f1z;
    const float d20 = qx * f0x + qy * f0y + qz * f0z, d21 = qx * f1x + qy * f1y + qz * f1z;
    const float den = d00 * d11 - d01 * d01;
    const float u = den != 0.0f ? (d11 * d20 - d01 * d21) / den : 0.0f;
    const float v = den != 0.0f ? (d00 * d21 - d01 * d20) / den : 0.0f;
    o->uv[0] = u;
    o->uv[1] = v;
    /* shading normal: u n0 + v n1 + (1 - u - v) n2, normalised */
    float nsx = ngx, nsy = ngy, nsz = ngz;
    if ((__float_as_uint(bb.w) & 1u) && b.tri_n) {
        const float4* tn = reinterpret_cast<const float4*>(b.tri_n) + 3 * (size_t)h.id;
        const float4 n0 = tn[0], n1 = tn[1], n2 = tn[2];
        const float w = 1.0f - u - v;
        const float sx = n0.x * u + n1.x * v + n2.x * w, sy = n0.y * u + n1.y * v + n2.y * w,
                    sz = n0.z * u + n1.z * v + n2.z * w;
        const float l2 = sx * sx + sy * sy + sz * sz;
        if (l2 > 0.0f) {
            const float il = 1.0f / sqrtf(l2);
            nsx = sx * il;
            nsy = sy * il;
            nsz = sz * il;
        }
    }
    o->shading_normal[0] = nsx;
    o->shading_normal[1] = nsy;
    o->shading_normal[2] = nsz;
}

/* Stores record r as entry i of the structure-of-arrays buffers `o` (the
 * reference's IntersectionBuffer layout, src/core/intersection.h:62-120);
 * NULL fields are skipped, and a miss writes only intersected and t. */
__device__ __forceinline__ void rb_isect_store(const rb_isect_soa& o, uint32_t i, const rb_isect_one& r) {
    if (o.intersected) o.intersected[i] = r.intersected;
    if (o.t) o.t[i] = r.t;
    if (!r.intersected) return;
    if (o.primitive_id) o.primitive_id[i] = (uint32_t)r.primitive_id;
    if (o.material_id) o.material_id[i] = r.material_id;
    for (int k = 0; k < 3; k++) {
        if (o.position[k]) o.position[k][i] = r.position[k];
        if (o.geometric_normal[k]) o.geometric_normal[k][i] = r.geometric_normal[k];
        if (o.shading_normal[k]) o.shading_normal[k][i] = r.shading_normal[k];
        if (o.incoming_direction[k]) o.incoming_direction[k][i] = r.incoming_direction[k];
    }
    if (o.front) o.front[i] = r.front;
    if (o.epsilon) o.epsilon[i] = r.epsilon;
    if (o.uv) {
        o.uv[2 * (size_t)i] = r.uv[0];
        o.uv[2 * (size_t)i + 1] = r.uv[1];
    }
}

/* trace_ray(ray, anyhit, isect, ...): traces the ray and, for a closest
 * hit, fills *isect (when not NULL); for any hit only intersected and t.
 * Returns whether the ray hit. */
__device__ __forceinline__ bool rb_trace_one_isect(const rb_bvh_dev& b, const rb_ray& ray, uint32_t flags,
                                                   rb_isect_one* isect) {
    const rb_hit h = rb_trace_one(b, ray, flags);
    if (isect) {
        if (flags & RB_ANY) {
            isect->intersected = h.id >= 0 ? 1 : 0;
            isect->t = h.t;
        } else {
            rb_fill_one(b, ray, h, isect);
        }
    }
    return h.id >= 0;
}

#endif /* RB_TRACE_DEVICE_H */
