// rb_trace_one.hip -- the C-ABI around the header-only device tracer
// (include/rb_trace_device.h): the device view of a resident BVH and a
// one-ray-per-thread kernel built on rb_trace_one().
#include <cstring>

#include "rb_device.h"
#include "../../include/rb_trace_device.h"

namespace rb {

__global__ void k_trace_simple(rb_bvh_dev b, const rb_ray* __restrict__ rays, uint32_t n, uint32_t flags,
                               rb_hit* __restrict__ hits) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        hits[i] = rb_trace_one(b, rays[i], flags);
}

// One ray per thread: trace, then fill and store the intersection record.
__global__ void k_trace_simple_fill(rb_bvh_dev b, const rb_ray* __restrict__ rays, uint32_t n, uint32_t flags,
                                    rb_isect_soa o) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        rb_isect_one rec;
        rb_trace_one_isect(b, rays[i], flags, &rec);
        if (flags & RB_ANY) {  // the reference fills nothing else for any hit
            if (o.intersected) o.intersected[i] = rec.intersected;
            if (o.t) o.t[i] = rec.t;
        } else {
            rb_isect_store(o, i, rec);
        }
    }
}

rb_bvh_dev device_view(rb_bvh_t bvh) {
    rb_bvh_dev v;
    std::memset(&v, 0, sizeof(v));
    v.nodes = bvh->nodes;
    v.tris = bvh->tris;
    v.index = bvh->index;
    v.n_nodes = bvh->n_nodes;
    v.n_refs = bvh->n_refs;
    v.errors = dev().errors;
    if (dev().cfg.wide && !dev().cfg.speculative && bvh->wnodes) {
        v.wnodes = bvh->wnodes;
        v.n_wnodes = bvh->n_wnodes;
    }
    v.n_tris = bvh->tri_v ? bvh->n_tris : 0;
    v.tri_v = bvh->tri_v;
    v.tri_n = bvh->tri_n;
    return v;
}

static uint32_t simple_grid(uint32_t n) {
    uint64_t g = ((uint64_t)n + 255) / 256;
    const uint64_t cap = (uint64_t)dev().num_cus * 32;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

}  // namespace rb

using namespace rb;

extern "C" {

rb_status rb_bvh_device_view(rb_bvh_t bvh, rb_bvh_dev* out) {
    if (!bvh || !out) return fail(RB_ERR_INVALID, "NULL argument");
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    *out = device_view(bvh);
    return RB_OK;
}

rb_status rb_trace_simple(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, rb_hit* hits, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (n == 0) return RB_OK;
    if (!bvh || !rays || !hits) return fail(RB_ERR_INVALID, "NULL bvh, rays or hits");
    if (flags & RB_MASKED) return fail(RB_ERR_INVALID, "rb_trace_simple takes no mask");
    if (flags & ~(uint32_t)RB_ANY) return fail(RB_ERR_INVALID, "unknown trace flags 0x%x", flags);
    hipLaunchKernelGGL(k_trace_simple, dim3(simple_grid(n)), dim3(256), 0, stream ? (hipStream_t)stream : dev().stream,
                       device_view(bvh), rays, n, flags & RB_ANY, hits);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RB_OK : hip_fail(e, "simple trace launch");
}

rb_status rb_trace_simple_fill(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, const rb_isect_soa* out,
                               void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (n == 0) return RB_OK;
    if (!bvh || !rays || !out) return fail(RB_ERR_INVALID, "NULL bvh, rays or output");
    if (flags & RB_MASKED) return fail(RB_ERR_INVALID, "rb_trace_simple_fill takes no mask");
    if (flags & ~(uint32_t)RB_ANY) return fail(RB_ERR_INVALID, "unknown trace flags 0x%x", flags);
    if (!bvh->tri_v) return fail(RB_ERR_INVALID, "rb_trace_simple_fill needs a BVH uploaded with its scene");
    hipLaunchKernelGGL(k_trace_simple_fill, dim3(simple_grid(n)), dim3(256), 0,
                       stream ? (hipStream_t)stream : dev().stream, device_view(bvh), rays, n, flags & RB_ANY, *out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RB_OK : hip_fail(e, "simple trace-and-fill launch");
}

}  // extern "C"
