// rb_trace_one.hip -- the C-ABI around the header-only device tracer
// (include/rb_trace_device.h): the device view of a resident BVH and a
// one-ray-per-thread kernel built on rb_trace_one().
#include "rb_device.h"
#include "../../include/rb_trace_device.h"

namespace rb {

__global__ void k_trace_simple(rb_bvh_dev b, const rb_ray* __restrict__ rays, uint32_t n, uint32_t flags,
                               rb_hit* __restrict__ hits) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        hits[i] = rb_trace_one(b, rays[i], flags);
}

static rb_bvh_dev view_of(rb_bvh_t bvh) {
    rb_bvh_dev v;
    v.nodes = bvh->nodes;
    v.tris = bvh->tris;
    v.index = bvh->index;
    v.n_nodes = bvh->n_nodes;
    v.n_refs = bvh->n_refs;
    v.errors = dev().errors;
    return v;
}

}  // namespace rb

using namespace rb;

extern "C" {

rb_status rb_bvh_device_view(rb_bvh_t bvh, rb_bvh_dev* out) {
    if (!bvh || !out) return fail(RB_ERR_INVALID, "NULL argument");
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    *out = view_of(bvh);
    return RB_OK;
}

rb_status rb_trace_simple(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, rb_hit* hits, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (n == 0) return RB_OK;
    if (!bvh || !rays || !hits) return fail(RB_ERR_INVALID, "NULL bvh, rays or hits");
    if (flags & RB_MASKED) return fail(RB_ERR_INVALID, "rb_trace_simple takes no mask");
    uint64_t g = ((uint64_t)n + 255) / 256;
    const uint64_t cap = (uint64_t)dev().num_cus * 32;
    if (g > cap) g = cap;
    hipLaunchKernelGGL(k_trace_simple, dim3((uint32_t)g), dim3(256), 0, stream ? (hipStream_t)stream : dev().stream,
                       view_of(bvh), rays, n, flags & RB_ANY, hits);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RB_OK : hip_fail(e, "simple trace launch");
}

}  // extern "C"
