// rb_device.hip -- device bookkeeping of the C-ABI: init, memory, streams,
// events, and the BVH/scene upload (CudaTracer ctor + Scene::gpu_prepare).
#include <mutex>

#include "rb_device.h"

namespace rb {

static DeviceState g_dev;
static std::mutex g_init_mu;
thread_local CaptureCtx* t_capture = nullptr;
thread_local Arena* t_arena = nullptr;

DeviceState& dev() { return g_dev; }

rb_status hip_fail(hipError_t e, const char* what) {
    return fail(RB_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

static rb_status init_locked(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(RB_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= n) return fail(RB_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    g_dev.num_cus = (uint32_t)prop.multiProcessorCount;
    if (prop.sharedMemPerBlock) g_dev.lds_per_block = prop.sharedMemPerBlock;
    if ((e = hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "hipStreamCreate");
    void* c = nullptr;
    const size_t cbytes = sizeof(uint32_t) * ((size_t)kCounterWords * kCounterSlots + kErrorWords);
    if ((e = hipMalloc(&c, cbytes)) != hipSuccess) return hip_fail(e, "hipMalloc(counters)");
    if ((e = hipMemset(c, 0, cbytes)) != hipSuccess) return hip_fail(e, "hipMemset(counters)");
    g_dev.counters = (uint32_t*)c;
    g_dev.errors = g_dev.counters + (size_t)kCounterWords * kCounterSlots;
    // the allocation-free entry point's scratch address; if the lookup
    // fails, that entry point reports it and everything else still works
    (void)noalloc_setup(device);
    (void)hipGetLastError();
    // Render scratch comes from the stream-ordered pool; keep freed blocks in
    // the pool instead of returning them at every synchronisation, so a
    // steady stream of renders reuses its memory rather than remapping it.
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
        uint64_t keep = ~0ull;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    // the best measured shape on 2 Mi incoherent rays at 1M triangles: the
    // 12-entry LDS stack and 384 cached top-of-tree records take 48 KB per
    // 512-thread block, which keeps 3 blocks (24 waves) on a CU
    g_dev.cfg.block = 512;
    g_dev.cfg.lds_stack = 12;
    g_dev.cfg.chunk = 64;
    g_dev.cfg.fetch_threshold = 60;
    g_dev.cfg.waves_per_cu = 24;
    g_dev.cfg.speculative = 0;
    g_dev.cfg.leaf_gate = 24;
    g_dev.cfg.lds_nodes = 384;
    g_dev.cfg.wide = 1;
    g_dev.device = device;
    return RB_OK;
}

rb_status ensure_device() {
    if (g_dev.device >= 0) {
        if (!t_capture) (void)hipSetDevice(g_dev.device);  // no device calls inside a graph capture
        return RB_OK;
    }
    std::lock_guard<std::mutex> g(g_init_mu);
    if (g_dev.device >= 0) return RB_OK;
    return init_locked(0);
}

// Per ring slot: the stream of its last launch and an event recorded after
// it (created on the slot's first use).  Striped locks keep a take and its
// release atomic against another thread wrapping round to the same slot.
static hipStream_t g_slot_stream[kCounterSlots];
static hipEvent_t g_slot_event[kCounterSlots];
static std::mutex g_slot_mu[64];

SlotLease take_counter_slot(hipStream_t stream) {
    SlotLease l;
    if (t_capture) {
        if (t_capture->next_pair < t_capture->n_pairs)
            l.counter = t_capture->counters + (size_t)kCounterWords * t_capture->next_pair++;
        return l;
    }
    const uint32_t s = g_dev.next_slot.fetch_add(1) % kCounterSlots;
    g_slot_mu[s % 64].lock();
    hipEvent_t& ev = g_slot_event[s];
    if (!ev) {
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            ev = nullptr;
            g_slot_mu[s % 64].unlock();
            l.st = hip_fail(e, "work-counter slot event");
            return l;
        }
    } else if (g_slot_stream[s] != stream) {
        hipError_t e = hipStreamWaitEvent(stream, ev, 0);
        if (e != hipSuccess) {
            g_slot_mu[s % 64].unlock();
            l.st = hip_fail(e, "work-counter slot wait");
            return l;
        }
    }
    l.counter = g_dev.counters + (size_t)kCounterWords * s;
    l.slot = (int32_t)s;
    return l;  // the slot's lock is held until release_counter_slot
}

void release_counter_slot(const SlotLease& l, hipStream_t stream) {
    if (l.slot < 0) return;
    const uint32_t s = (uint32_t)l.slot;
    if (hipEventRecord(g_slot_event[s], stream) == hipSuccess) {
        g_slot_stream[s] = stream;
    } else {
        // no event marks this launch's end: finish it before the slot is
        // handed out again (error path only)
        (void)hipGetLastError();
        (void)hipStreamSynchronize(stream);
        (void)hipGetLastError();
    }
    g_slot_mu[s % 64].unlock();
}

rb_status dalloc(void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return RB_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? RB_ERR_OOM : RB_ERR_DEVICE, "hipMalloc(%zu): %s", bytes,
                    hipGetErrorString(e));
    }
    g_dev.library_bytes += bytes;
    return RB_OK;
}

void dfree(void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    g_dev.library_bytes -= bytes;
}

}  // namespace rb

using namespace rb;

static hipStream_t S(void* s) { return s ? (hipStream_t)s : dev().stream; }

#define RB_ENSURE()                               \
    do {                                          \
        rb_status _st = ensure_device();          \
        if (_st != RB_OK) return _st;             \
    } while (0)

#define RB_HIP(call, what)                         \
    do {                                           \
        hipError_t _e = (call);                    \
        if (_e != hipSuccess) return hip_fail(_e, what); \
    } while (0)

extern "C" {

rb_status rb_device_init(int device) {
    std::lock_guard<std::mutex> g(g_init_mu);
    if (g_dev.device >= 0) {
        if (g_dev.device == device) return RB_OK;
        return fail(RB_ERR_BUSY, "library already bound to device %d (one GPU per process)", g_dev.device);
    }
    return init_locked(device);
}

rb_status rb_device_name(char* buf, size_t n, uint32_t* num_cus) {
    RB_ENSURE();
    hipDeviceProp_t prop;
    RB_HIP(hipGetDeviceProperties(&prop, dev().device), "hipGetDeviceProperties");
    if (buf && n) {
        snprintf(buf, n, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (num_cus) *num_cus = dev().num_cus;
    return RB_OK;
}

rb_status rb_malloc(void** ptr, size_t bytes) {
    if (!ptr) return fail(RB_ERR_INVALID, "ptr is NULL");
    RB_ENSURE();
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return fail(e == hipErrorOutOfMemory ? RB_ERR_OOM : RB_ERR_DEVICE, "hipMalloc(%zu): %s", bytes,
                    hipGetErrorString(e));
    }
    return RB_OK;
}

rb_status rb_free(void* ptr) {
    if (!ptr) return RB_OK;
    RB_ENSURE();
    RB_HIP(hipFree(ptr), "hipFree");
    return RB_OK;
}

rb_status rb_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
    RB_ENSURE();
    if (!bytes) return RB_OK;
    RB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)), "hipMemcpy H2D");
    RB_HIP(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
    return RB_OK;
}

rb_status rb_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
    RB_ENSURE();
    if (!bytes) return RB_OK;
    RB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)), "hipMemcpy D2H");
    RB_HIP(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
    return RB_OK;
}

rb_status rb_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
    RB_ENSURE();
    if (!bytes) return RB_OK;
    RB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)), "hipMemcpy D2D");
    return RB_OK;
}

rb_status rb_memset(void* dst, int value, size_t bytes, void* stream) {
    RB_ENSURE();
    if (!bytes) return RB_OK;
    RB_HIP(hipMemsetAsync(dst, value, bytes, S(stream)), "hipMemsetAsync");
    return RB_OK;
}

}  // extern "C"

namespace rb {
__global__ void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace rb

extern "C" {

rb_status rb_copy_kernel(void* dst, const void* src, size_t bytes, void* stream) {
    RB_ENSURE();
    if (bytes == 0) return RB_OK;
    if (!dst || !src || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return fail(RB_ERR_INVALID, "rb_copy_kernel needs 16-byte aligned pointers and sizes");
    const size_t n16 = bytes / 16;
    size_t grid = (n16 + 255) / 256;
    const size_t cap = (size_t)dev().num_cus * 32;
    if (grid > cap) grid = cap;
    hipLaunchKernelGGL(k_copy16, dim3((uint32_t)grid), dim3(256), 0, S(stream), (uint4*)dst, (const uint4*)src, n16);
    RB_HIP(hipGetLastError(), "copy kernel launch");
    return RB_OK;
}

rb_status rb_mem_info(size_t* free_bytes, size_t* total_bytes, size_t* library_bytes) {
    RB_ENSURE();
    size_t f = 0, t = 0;
    RB_HIP(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    if (library_bytes) *library_bytes = dev().library_bytes.load();
    return RB_OK;
}

rb_status rb_stream_create(void** stream) {
    if (!stream) return fail(RB_ERR_INVALID, "stream is NULL");
    RB_ENSURE();
    hipStream_t s;
    RB_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    *stream = s;
    return RB_OK;
}

rb_status rb_stream_destroy(void* stream) {
    if (!stream) return RB_OK;
    RB_ENSURE();
    RB_HIP(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
    return RB_OK;
}

rb_status rb_stream_synchronize(void* stream) {
    RB_ENSURE();
    RB_HIP(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
    return RB_OK;
}

rb_status rb_event_create(void** ev) {
    if (!ev) return fail(RB_ERR_INVALID, "ev is NULL");
    RB_ENSURE();
    hipEvent_t e;
    RB_HIP(hipEventCreate(&e), "hipEventCreate");
    *ev = e;
    return RB_OK;
}

rb_status rb_event_destroy(void* ev) {
    if (!ev) return RB_OK;
    RB_ENSURE();
    RB_HIP(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy");
    return RB_OK;
}

rb_status rb_event_record(void* ev, void* stream) {
    RB_ENSURE();
    RB_HIP(hipEventRecord((hipEvent_t)ev, S(stream)), "hipEventRecord");
    return RB_OK;
}

rb_status rb_event_elapsed_ms(void* a, void* b, float* ms) {
    RB_ENSURE();
    RB_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b), "hipEventElapsedTime");
    return RB_OK;
}

rb_status rb_event_synchronize(void* ev) {
    RB_ENSURE();
    RB_HIP(hipEventSynchronize((hipEvent_t)ev), "hipEventSynchronize");
    return RB_OK;
}

// ------------------------------------------------------------- upload --

rb_status rb_bvh_upload(rb_host_bvh_t hb, rb_scene_t scene, rb_bvh_t* out) {
    if (!hb || !out) return fail(RB_ERR_INVALID, "NULL argument");
    RB_ENSURE();
    Mi355Layout L;
    std::vector<uint32_t> wide;
    try {
        rb_status st = export_mi355x(*hb->b, L);
        if (st != RB_OK) return st;
        // every stack entry a ray can need must exist: the kernels hold kMaxStack
        const uint32_t depth = layout_depth(L.nodes);
        if (depth == UINT32_MAX || depth + 1 >= (uint32_t)kMaxStack)
            return fail(RB_ERR_UNSUPPORTED, "BVH node layout is %s (the traversal stack holds %d entries)",
                        depth == UINT32_MAX ? "not a tree" : "too deep", kMaxStack);
        // wide records only where their walk fits the same stack and their
        // 8-bit steps can hold the boxes (else the binary walk alone)
        if (wide_fits_stack(depth)) {
            st = export_wide(L.nodes, wide);
            if (st == RB_ERR_UNSUPPORTED) {
                wide.clear();
                set_error("");
            } else if (st != RB_OK) {
                return st;
            }
        }
    } catch (...) {
        return fail(RB_ERR_OOM, "layout export failed (host allocation)");
    }
    if (scene && scene->s.num_tris() != (uint32_t)(hb->b->idx.size() / 3))
        return fail(RB_ERR_INVALID, "scene has %u triangles, the BVH was built over %zu", scene->s.num_tris(),
                    hb->b->idx.size() / 3);
    std::unique_ptr<rb_bvh_s> B(new rb_bvh_s());
    auto up = [&](void** dst, const void* src, size_t bytes) -> rb_status {
        rb_status st = dalloc(dst, bytes);
        if (st != RB_OK) return st;
        B->bytes += bytes;
        if (bytes) {
            hipError_t e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(e, "hipMemcpy(upload)");
        }
        return RB_OK;
    };
    auto cleanup = [&]() { rb_bvh_release(B.release()); };
    rb_status st;
    B->n_nodes = (uint32_t)(L.nodes.size() / 16);
    B->n_refs = (uint32_t)L.index.size();
    if ((st = up((void**)&B->nodes, L.nodes.data(), L.nodes.size() * 4)) != RB_OK) { cleanup(); return st; }
    B->n_wnodes = (uint32_t)(wide.size() / kWideWords);
    if (!wide.empty() && (st = up((void**)&B->wnodes, wide.data(), wide.size() * 4)) != RB_OK) { cleanup(); return st; }
    // one spare row after the triangle array keeps 16-byte loads in bounds
    L.tris.resize(L.tris.size() + 4, 0.0f);
    if ((st = up((void**)&B->tris, L.tris.data(), L.tris.size() * 4)) != RB_OK) { cleanup(); return st; }
    L.index.push_back(0);
    if ((st = up((void**)&B->index, L.index.data(), L.index.size() * 4)) != RB_OK) { cleanup(); return st; }
    if (scene) {
        const Scene& Sc = scene->s;
        B->n_tris = Sc.num_tris();
        std::vector<float> tv((size_t)B->n_tris * 12);
        // shading normals: 3 float4 per triangle, only when some triangle has them
        std::vector<float> tn(Sc.nidx.empty() ? 0 : (size_t)B->n_tris * 12, 0.0f);
        for (uint32_t t = 0; t < B->n_tris; t++) {
            const bool has_n = Sc.tri_has_normals(t);
            for (int k = 0; k < 3; k++) {
                uint32_t vi = Sc.idx[3 * (size_t)t + k];
                tv[12 * (size_t)t + 4 * k + 0] = Sc.verts[3 * (size_t)vi];
                tv[12 * (size_t)t + 4 * k + 1] = Sc.verts[3 * (size_t)vi + 1];
                tv[12 * (size_t)t + 4 * k + 2] = Sc.verts[3 * (size_t)vi + 2];
                // v0.w: material; v1.w: flags (kTriHasNormals = 1)
                uint32_t w = k == 0 ? Sc.tri_mat[t] : (k == 1 && has_n ? 1u : 0u);
                std::memcpy(&tv[12 * (size_t)t + 4 * k + 3], &w, 4);
                if (has_n) {
                    const uint32_t ni = Sc.nidx[3 * (size_t)t + k];
                    for (int c = 0; c < 3; c++) tn[12 * (size_t)t + 4 * k + c] = Sc.normals[3 * (size_t)ni + c];
                }
            }
        }
        if ((st = up((void**)&B->tri_v, tv.data(), tv.size() * 4)) != RB_OK) { cleanup(); return st; }
        if (!tn.empty() && (st = up((void**)&B->tri_n, tn.data(), tn.size() * 4)) != RB_OK) { cleanup(); return st; }
        B->n_mats = (uint32_t)Sc.mats.size();
        if ((st = up((void**)&B->mats, Sc.mats.data(), Sc.mats.size() * sizeof(rb_material))) != RB_OK) { cleanup(); return st; }
        B->n_lights = (uint32_t)Sc.light_tris.size();
        if ((st = up((void**)&B->light_tris, Sc.light_tris.data(), Sc.light_tris.size() * 4)) != RB_OK) { cleanup(); return st; }
        float lo[3], hi[3];
        rb_scene_bounds(scene, lo, hi);
        for (int i = 0; i < 3; i++) { B->scene_lo[i] = lo[i]; B->scene_hi[i] = hi[i]; }
    }
    *out = B.release();
    return RB_OK;
}

rb_status rb_bvh_release(rb_bvh_t b) {
    if (!b) return RB_OK;
    if (dev().device >= 0) {
        dfree(b->nodes, (size_t)b->n_nodes * 64);
        dfree(b->wnodes, (size_t)b->n_wnodes * kWideWords * 4);
        dfree(b->tris, ((size_t)b->n_refs * 12 + 4) * 4);
        dfree(b->index, ((size_t)b->n_refs + 1) * 4);
        dfree(b->tri_v, (size_t)b->n_tris * 48);
        dfree(b->tri_n, (size_t)b->n_tris * 48);
        dfree(b->mats, (size_t)b->n_mats * sizeof(rb_material));
        dfree(b->light_tris, (size_t)b->n_lights * 4);
    }
    delete b;
    return RB_OK;
}

rb_status rb_bvh_device_bytes(rb_bvh_t b, size_t* bytes) {
    if (!b || !bytes) return fail(RB_ERR_INVALID, "NULL argument");
    *bytes = b->bytes;
    return RB_OK;
}

}  // extern "C"
