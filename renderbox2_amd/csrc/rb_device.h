// rb_device.h -- device-side structures shared by the HIP sources.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>

#include "rb_internal.h"
#include "rb_rng.h"

namespace rb {

// Work-counter slots: each launch of a persistent kernel takes one slot, a
// block of kCounterWords words: [1] the exit counter, [kBandStride * (b + 1)]
// the work counter of ray band b (b < kRayBands), each on a 128-byte line of
// its own (with all eight on one line, claims from different XCDs contend
// for it: 1.43 against 0.97 ms on a 2 Mi-ray launch); [0] the one work
// counter of a batched launch; the rest unused.  The
// last block out resets them to zero, so no memset node is needed between
// launches or inside captured graphs.  A
// slot is reused kCounterSlots launches later; when that launch is on
// another stream it first waits (stream-ordered, no host wait) for the
// event recorded after the slot's previous launch, so two launches never
// share a counter whatever the streams and the number in flight.
constexpr uint32_t kCounterSlots = 4096;
// A launch's rays are split into this many contiguous bands, one per XCD:
// the waves of an XCD claim rays from the band of its number first, so the
// rays that share an XCD's L2 start near each other (contiguous rays are
// neighbouring pixels' rays), then from the other bands in turn.
constexpr uint32_t kRayBands = 8;
constexpr uint32_t kBandStride = 32;  // words: one 128-byte line per counter
constexpr uint32_t kCounterWords = kBandStride * (kRayBands + 1);
// Words after the counter blocks: [0] traversal error flags (RB_TRACE_ERR_*).
constexpr uint32_t kErrorWords = 64;

// Traversal stack bound: the builder stops at depth 64 and leaves of more than
// 31 references are split by the exporter, so 96 entries never overflow.
constexpr int kMaxStack = (int)kTraceStackEntries;
constexpr int kSentinel = 0x76543210;

struct DeviceState {
    int device = -1;
    uint32_t num_cus = 0;
    size_t lds_per_block = 65536;  // LDS one workgroup may use (static + dynamic)
    hipStream_t stream = nullptr;  // library default stream
    uint32_t* counters = nullptr;  // kCounterWords * kCounterSlots, then kErrorWords
    uint32_t* errors = nullptr;    // counters + kCounterWords * kCounterSlots
    std::atomic<uint32_t> next_slot{0};
    std::atomic<size_t> library_bytes{0};
    rb_trace_config cfg{};
};
DeviceState& dev();
rb_status ensure_device();
// Records the device address of the allocation-free entry point's static
// scratch array for `device` (rb_raygen.hip); run once per device by the
// init, with that device current.
rb_status noalloc_setup(int device);
// A launch's work-counter block: from the ring, or from the pool of the render
// graph being captured on this thread (counter == nullptr when that pool is
// used up).  take_counter_slot makes `stream` wait for the slot's previous
// launch when that was on another stream; release_counter_slot, called once
// the launch is enqueued, records the slot's event.  Every take is followed
// by exactly one release.
struct SlotLease {
    uint32_t* counter = nullptr;
    int32_t slot = -1;  // ring slot, -1 for a graph's own pair
    rb_status st = RB_OK;
};
SlotLease take_counter_slot(hipStream_t stream);
void release_counter_slot(const SlotLease& l, hipStream_t stream);

// Set on the capturing thread while a render pass is recorded into a graph
// (rb_render_graph_create): kernels read their stream keys from `keys`
// (indexed by tag), trace launches take counter blocks from the graph's own
// pool, and scratch comes from persistent allocations the graph frees.
struct CaptureCtx {
    const Rng* keys = nullptr;
    uint32_t* counters = nullptr;
    uint32_t n_pairs = 0, next_pair = 0;
    std::vector<void*> allocs;
};
extern thread_local CaptureCtx* t_capture;

// Set on the server's scheduler thread while it launches a request: render
// scratch is carved from the slot's own device arena, which only that slot's
// stream touches, so requests on different streams never reuse each other's
// memory (no allocator call, no cross-stream ordering).  Allocations that do
// not fit fall back to the stream-ordered pool.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, off = 0;
    uint32_t misses = 0;  // allocations that did not fit
    void* take(size_t bytes) {
        const size_t need = (bytes + 255) & ~(size_t)255;
        if (!base || off + need > cap) {
            misses++;
            return nullptr;
        }
        void* p = base + off;
        off += need;
        return p;
    }
};
extern thread_local Arena* t_arena;
inline RngArg rng_arg(uint64_t seed, uint32_t tag) {
    RngArg a;
    a.r = Rng::make(seed, tag);
    a.dev = t_capture && t_capture->keys ? t_capture->keys + tag : nullptr;
    return a;
}
rb_status hip_fail(hipError_t e, const char* what);

// Device view of a resident BVH for rb_trace_one (rb_trace_one.hip): its
// records, the wide ones when the traversal config walks them, and the
// scene rows it was uploaded with.
rb_bvh_dev device_view(rb_bvh_t bvh);

// Internal trace flag, beside RB_ANY: an any-hit set whose consumer reads
// only hit / miss (hit.x < 0).  Such a set may walk the wide records, which
// find a hit exactly when the binary walk does but not necessarily the same
// first one.  The public entry points refuse the bit.
constexpr uint32_t kOcclusion = 1u << 16;

// Persistent traversal launch (rb_trace.hip).  n_dev, when set, holds the ray
// count in device memory (queues sized by earlier kernels); n is then an
// upper bound that only sizes the grid.  counter, when set, is the launch's
// zeroed work-counter block (kCounterWords words the kernel leaves zero again) instead
// of a slot of the ring.
rb_status launch_trace(rb_bvh_t bvh, const rb_ray* rays, uint32_t n, uint32_t flags, const uint8_t* mask,
                       rb_hit* hits, uint32_t* counts, hipStream_t stream, const rb_trace_config& cfg,
                       const uint32_t* n_dev = nullptr, uint32_t* counter = nullptr);
// Two queued ray sets of one BVH in one persistent launch: set 0 any hit
// (flags0 = RB_ANY) or closest hit (RB_CLOSEST), set 1 closest hit, their
// counts in device memory (cap0 / cap1 bound them and size the grid).  Same
// per-ray results as two launch_trace calls, with one launch's ramp-up and
// tail instead of two.
rb_status launch_trace_pair(rb_bvh_t bvh, const rb_ray* rays0, uint32_t cap0, const uint32_t* n0_dev, rb_hit* hits0,
                            uint32_t flags0, const rb_ray* rays1, uint32_t cap1, const uint32_t* n1_dev, rb_hit* hits1,
                            hipStream_t stream, const rb_trace_config& cfg);

// A ray set of one BVH for a batched launch: n rays, or *n_dev of them
// (bounded by n) when the count is in device memory.
struct BatchSet {
    rb_bvh_t bvh;
    const float4* rays;
    int2* hits;
    uint32_t n;
    const uint32_t* n_dev;
    bool any;
    const uint8_t* mask;  // NULL: every ray is traced
    bool occlusion = false;  // any hit read only as hit / miss (kOcclusion)
};
// Ray sets of any resident BVHs through batched persistent launches (up to
// 64 sets and 2^31 - 1 rays of bound per launch).  Same per-ray results as
// one launch_trace per set.
rb_status launch_trace_sets(const BatchSet* sets, uint32_t ns, hipStream_t stream, const rb_trace_config& cfg);

// A ray set a render step leaves to trace.
struct TraceSet {
    const float4* rays;
    int2* hits;
    uint32_t n;             // rays, or the bound of *n_dev
    const uint32_t* n_dev;  // device count (NULL: n)
    uint32_t flags;         // RB_CLOSEST or RB_ANY (with kOcclusion), with RB_MASKED
    const uint8_t* mask;
    rb_trace_config cfg;    // launch shape when the set is traced on its own
};

// A render call as a sequence of steps (rb_render.hip).  step() enqueues on
// s the kernels up to the job's next trace and returns that trace's ray sets
// (one or two); *done once the last step has written the film.  All device
// memory is taken when the job is made (from the server's arena when
// t_arena is set) and released when it is destroyed.
class RenderJob {
  public:
    virtual ~RenderJob() {}
    virtual rb_status step(hipStream_t s, TraceSet* sets, uint32_t* nsets, bool* done) = 0;
    rb_bvh_t bvh = nullptr;
};
// The job of one render call of kind RB_REQ_AO / _PT / _BDPT / _RAYCAST
// (the checks of the render_* calls; s: the stream its memory is freed on).
rb_status make_render_job(uint32_t kind, rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w,
                          uint32_t h, const rb_render_params* p, float* out, hipStream_t s,
                          std::unique_ptr<RenderJob>* job);
// Runs a job on one stream, each step's sets in one launch.
rb_status run_job(RenderJob& J, hipStream_t s);

// Render passes (rb_render.hip), asynchronous on stream s.
rb_status render_ao(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                    const rb_render_params* p, float* out, hipStream_t s);
rb_status render_pt(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                    const rb_render_params* p, float* out, hipStream_t s);
rb_status render_bdpt(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                      const rb_render_params* p, float* out, hipStream_t s);
rb_status render_raycast(rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                         const rb_render_params* p, float* out, hipStream_t s);
// Device scratch bytes one render call of this kind and size allocates.
uint64_t render_scratch_bytes(uint32_t kind, uint32_t w, uint32_t h, const rb_render_params* p);

// Device allocation tracked against the library's HBM accounting
// (MemoryAllocator::allocate, src/memory/memoryallocator.cpp:54-67).
rb_status dalloc(void** p, size_t bytes);
void dfree(void* p, size_t bytes);

}  // namespace rb

// Device-resident BVH + optional shading data.
struct rb_bvh_s {
    float4* nodes = nullptr;    // 4 float4 per inner node (MI355X layout)
    float4* wnodes = nullptr;   // 4 float4 per wide (4-child) record (export_wide)
    uint32_t n_wnodes = 0;
    float4* tris = nullptr;     // 3 float4 Woop rows per triangle reference
    uint32_t* index = nullptr;  // original triangle id per reference
    uint32_t n_nodes = 0, n_refs = 0;
    // shading data (Scene::gpu_prepare, src/core/scene.cpp:166-301)
    float4* tri_v = nullptr;          // 3 float4 per triangle: v0.xyz|mat, v1.xyz|flags, v2.xyz|0
    float4* tri_n = nullptr;          // 3 float4 per triangle (n0, n1, n2) when the scene has normals
    rb_material* mats = nullptr;
    uint32_t* light_tris = nullptr;
    uint32_t n_tris = 0, n_mats = 0, n_lights = 0;
    float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {0, 0, 0};
    size_t bytes = 0;
};
