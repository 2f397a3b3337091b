// rb_raygen.hip -- ray generation and the intersection-fill pass.
//  * camera rays: kernel_generate_perspective_camera_samples
//    (renderbox2 src/renderers/samplerrenderer.cu:48-96) + generate_ray
//    (src/core/camera.cu:91-98), Philox-keyed per (pixel, sample);
//  * incoherent diffuse rays from primary hits (the trace benchmark's
//    second ray set);
//  * intersection fill: the result-write block of kernel_kepler_trace
//    (src/accelerators/sbvh/kepler_dynamic_fetch.h:312-360) as its own pass.
#include "rb_device.h"
#include "rb_sample.h"
#include "rb_shade.h"
#include "../../include/rb_trace_device.h"

namespace rb {

struct CamArgs {
    float r2c[16];
    float c2w[16];
};

// Matrix4x4::operator()(float3, point) (src/core/transform.h:182-197):
// float4 dot products summed left to right, then divide by w for points.
__device__ __forceinline__ V3 xform_point(const float* m, float x, float y, float z) {
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; i++) r[i] = m[4 * i] * x + m[4 * i + 1] * y + m[4 * i + 2] * z + m[4 * i + 3] * 1.0f;
    return v3(r[0] / r[3], r[1] / r[3], r[2] / r[3]);
}
__device__ __forceinline__ V3 xform_vector(const float* m, V3 v) {
    float r[3];
#pragma unroll
    for (int i = 0; i < 3; i++) r[i] = m[4 * i] * v.x + m[4 * i + 1] * v.y + m[4 * i + 2] * v.z + m[4 * i + 3] * 0.0f;
    return v3(r[0], r[1], r[2]);
}

__device__ __forceinline__ void camera_ray(const CamArgs& c, float px, float py, float4* out) {
    const V3 pc = xform_point(c.r2c, px, py, 0.0f);
    const V3 d = vnorm(pc);
    const V3 o = xform_point(c.c2w, 0.0f, 0.0f, 0.0f);
    const V3 dw = xform_vector(c.c2w, d);
    out[0] = make_float4(o.x, o.y, o.z, 0.0f);
    out[1] = make_float4(dw.x, dw.y, dw.z, 3.402823466e+38f);
}

__global__ void k_raygen_camera(CamArgs c, uint32_t width, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                                uint32_t spp, RngArg rng_a, int jitter, float4* __restrict__ rays) {
    const Rng rng = rng_a.get();
    const uint64_t total = (uint64_t)w * h * spp;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = (uint32_t)(i % spp);
        const uint32_t p = (uint32_t)(i / spp);
        const uint32_t px = x0 + p % w, py = y0 + p / w;
        float jx = 0.5f, jy = 0.5f;
        if (jitter) {
            const U4 r = rng.draw(py * width + px, s, 0, 0);
            jx = u01(r.x);
            jy = u01(r.y);
        }
        camera_ray(c, (float)px + jx, (float)py + jy, rays + 2 * i);
    }
}

__global__ void k_raygen_diffuse(const float4* __restrict__ prim, const int2* __restrict__ hits, uint32_t np,
                                 uint32_t nout, const float4* __restrict__ tri_v, Rng rng, Rng redraw,
                                 float4* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nout; i += gridDim.x * blockDim.x) {
        uint32_t j = i % np;
        int2 h = hits[j];
        for (uint32_t att = 0; h.x < 0 && att < 64; att++) {
            j = redraw.draw(i, att, 0, 0).x % np;
            h = hits[j];
        }
        if (h.x < 0) {  // nothing was hit anywhere near: an empty ray
            out[2 * (size_t)i] = make_float4(0.f, 0.f, 0.f, 0.f);
            out[2 * (size_t)i + 1] = make_float4(0.f, 0.f, 1.f, 0.f);
            continue;
        }
        const float4 o = prim[2 * (size_t)j], d = prim[2 * (size_t)j + 1];
        const float t = __int_as_float(h.y);
        const V3 p = vadd(v3(o.x, o.y, o.z), vmul(v3(d.x, d.y, d.z), t));
        const float4 a = tri_v[3 * (size_t)h.x], b = tri_v[3 * (size_t)h.x + 1], c = tri_v[3 * (size_t)h.x + 2];
        V3 n = tri_normal(v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), v3(c.x, c.y, c.z));
        if (vdot(n, v3(d.x, d.y, d.z)) > 0.0f) n = vneg(n);
        const U4 r = rng.draw(i, 0, 0, 0);
        const V3 dir = cosine_hemisphere(n, u01(r.x), u01(r.y));
        out[2 * (size_t)i] = make_float4(p.x, p.y, p.z, 1.0e-3f * t);
        out[2 * (size_t)i + 1] = make_float4(dir.x, dir.y, dir.z, 3.402823466e+38f);
    }
}

// The intersection records of a trace's hits: rb_fill_one of the public
// device header (the same arithmetic as a kernel that traces and fills in
// its own code), stored into the caller's arrays.
__global__ void k_fill(const float4* __restrict__ rays, const int2* __restrict__ hits, uint32_t n, SceneDev S,
                       rb_isect_soa o) {
    rb_bvh_dev b;
    memset(&b, 0, sizeof(b));  // only the scene rows are read
    b.tri_v = S.tri_v;
    b.tri_n = S.tri_n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int2 h2 = hits[i];
        const float4 ro = rays[2 * (size_t)i], rd = rays[2 * (size_t)i + 1];
        rb_ray r;
        r.ox = ro.x; r.oy = ro.y; r.oz = ro.z; r.tmin = ro.w;
        r.dx = rd.x; r.dy = rd.y; r.dz = rd.z; r.tmax = rd.w;
        rb_hit h;
        h.id = h2.x;
        h.t = __int_as_float(h2.y);
        rb_isect_one rec;
        rb_fill_one(b, r, h, &rec);
        rb_isect_store(o, i, rec);
    }
}

static uint32_t grid_for(uint64_t n, uint32_t block) {
    uint64_t g = (n + block - 1) / block;
    const uint64_t cap = (uint64_t)dev().num_cus * 16;
    if (g > cap) g = cap;
    return (uint32_t)(g ? g : 1);
}

// ------------------------------------------ allocation-free trace + fill --
// rb_trace_fill_noalloc keeps what rb_trace_fill would allocate -- the
// (id, t) staging between the trace and the fill, and the trace's
// work-counter block -- in this file-scope device array: bytes
// [0, kScratchHead) hold the counter block, then kNoAllocRays rb_hit
// records.  Each device holds its own copy for the life of the process.  In
// host code `scratch` names a host-side stand-in; kernels get its device
// address, looked up once per device at init (noalloc_setup).
constexpr uint32_t kNoAllocRays = 1u << 20;
constexpr size_t kScratchHead = 1280;  // the counter block, rounded up to 256 bytes
static_assert(kScratchHead >= kCounterWords * sizeof(uint32_t), "the scratch head holds a counter block");
static __device__ __align__(256) unsigned char scratch[kScratchHead + (size_t)kNoAllocRays * sizeof(rb_hit)];
constexpr int kMaxDevices = 64;
static void* g_scratch_dev[kMaxDevices];  // device address of `scratch`, per device

__global__ void k_zero_words(uint32_t* p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// Trace into `hits`, then write the intersection records from them.
// `counter`: the trace's work-counter block when the caller brings one (else
// a slot of the library's ring).
static rb_status launch_fill(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, const uint8_t* mask,
                             rb_hit* hits, const rb_isect_soa* out, hipStream_t s, uint32_t* counter) {
    rb_status st = launch_trace(bvh, rays, n, flags, mask, hits, nullptr, s, dev().cfg, nullptr, counter);
    if (st != RB_OK) return st;
    SceneDev S;
    std::memset(&S, 0, sizeof(S));
    S.tri_v = bvh->tri_v;
    S.tri_n = bvh->tri_n;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256)), dim3(256), 0, s, reinterpret_cast<const float4*>(rays),
                       reinterpret_cast<const int2*>(hits), n, S, *out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? RB_OK : hip_fail(e, "fill launch");
}

rb_status noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(RB_ERR_INVALID, "device %d: no scratch slot", device);
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[device] = p;
    return RB_OK;
}

}  // namespace rb

using namespace rb;

extern "C" {

rb_status rb_raygen_camera(const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t spp,
                           uint64_t seed, int jitter, rb_ray* rays, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (!cam || !rays) return fail(RB_ERR_INVALID, "NULL camera or rays");
    if (spp == 0 || w == 0 || h == 0) return RB_OK;
    if ((uint64_t)x0 + w > cam->width || (uint64_t)y0 + h > cam->height)
        return fail(RB_ERR_INVALID, "tile outside the film");
    CamArgs c;
    if ((st = rb_camera_matrices(cam, c.r2c, c.c2w)) != RB_OK) return st;
    const uint64_t total = (uint64_t)w * h * spp;
    hipStream_t s = stream ? (hipStream_t)stream : dev().stream;
    hipLaunchKernelGGL(k_raygen_camera, dim3(grid_for(total, 256)), dim3(256), 0, s, c, cam->width, x0, y0, w, h, spp,
                       rng_arg(seed, RNG_CAMERA), jitter, reinterpret_cast<float4*>(rays));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RB_OK : hip_fail(e, "raygen launch");
}

rb_status rb_raygen_diffuse(rb_bvh_t bvh, const rb_ray* prim, const rb_hit* hits, uint32_t np, uint32_t nout,
                            uint64_t seed, rb_ray* out, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (!bvh || !prim || !hits || !out) return fail(RB_ERR_INVALID, "NULL argument");
    if (!bvh->tri_v) return fail(RB_ERR_INVALID, "bvh was uploaded without its scene");
    if (nout == 0) return RB_OK;
    if (np == 0) return fail(RB_ERR_INVALID, "no primary rays");
    hipStream_t s = stream ? (hipStream_t)stream : dev().stream;
    hipLaunchKernelGGL(k_raygen_diffuse, dim3(grid_for(nout, 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(prim), reinterpret_cast<const int2*>(hits), np, nout, bvh->tri_v,
                       Rng::make(seed, RNG_DIFFUSE), Rng::make(seed, RNG_REDRAW), reinterpret_cast<float4*>(out));
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RB_OK : hip_fail(e, "diffuse raygen launch");
}

rb_status rb_trace_fill(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, const uint8_t* mask,
                        rb_hit* hits, const rb_isect_soa* out, void* stream) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (!bvh || !out) return fail(RB_ERR_INVALID, "NULL argument");
    if (!bvh->tri_v) return fail(RB_ERR_INVALID, "bvh was uploaded without its scene");
    if (n == 0) return RB_OK;
    hipStream_t s = stream ? (hipStream_t)stream : dev().stream;
    // Without a caller buffer the (id, t) pairs live in a stream-ordered
    // allocation freed on the same stream after the fill: no host wait.
    rb_hit* tmp = nullptr;
    if (!hits) {
        hipError_t e = hipMallocAsync((void**)&tmp, (size_t)n * sizeof(rb_hit), s);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(e == hipErrorOutOfMemory ? RB_ERR_OOM : RB_ERR_DEVICE, "trace-fill hits (%zu bytes): %s",
                        (size_t)n * sizeof(rb_hit), hipGetErrorString(e));
        }
        hits = tmp;
    }
    st = launch_fill(bvh, rays, n, flags, mask, hits, out, s, nullptr);
    if (tmp) {
        const hipError_t e = hipFreeAsync(tmp, s);
        if (e != hipSuccess && st == RB_OK) st = hip_fail(e, "trace-fill hits free");
    }
    return st;
}

rb_status rb_trace_fill_noalloc(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, const uint8_t* mask,
                                const rb_isect_soa* out, void* stream) {
    // Nothing here allocates: the device must already be initialised (the
    // init allocates), and every buffer is the caller's or `scratch`.
    const int d = dev().device;
    if (d < 0) return fail(RB_ERR_INVALID, "rb_trace_fill_noalloc: call rb_device_init first");
    if (d >= kMaxDevices || !g_scratch_dev[d]) return fail(RB_ERR_DEVICE, "rb_trace_fill_noalloc: no scratch on device %d", d);
    if (!bvh || !rays || !out) return fail(RB_ERR_INVALID, "NULL argument");
    if (!bvh->tri_v) return fail(RB_ERR_INVALID, "bvh was uploaded without its scene");
    if (n > kNoAllocRays)
        return fail(RB_ERR_INVALID, "rb_trace_fill_noalloc: %u rays, at most %u per call (its scratch holds that many)",
                    n, kNoAllocRays);
    if (n == 0) return RB_OK;
    hipError_t e = hipSetDevice(d);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const hipStream_t s = stream ? (hipStream_t)stream : dev().stream;
    uint32_t* counter = static_cast<uint32_t*>(g_scratch_dev[d]);
    rb_hit* hits = reinterpret_cast<rb_hit*>(static_cast<char*>(g_scratch_dev[d]) + kScratchHead);
    // the counter block must start at zero (a kernel, not a memset: callers
    // may capture this into a graph)
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, counter, kCounterWords);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "scratch clear");
    return launch_fill(bvh, rays, n, flags, mask, hits, out, s, counter);
}

}  // extern "C"
