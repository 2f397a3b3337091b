// rb_graph.hip -- render passes captured into HIP graphs.
//
// A pass's bounce loop is a fixed chain of launches whose kernels read their
// queue lengths from device memory (no host synchronisation inside), so the
// whole pass -- camera rays, every bounce's shade / trace / connect /
// compaction, film -- is recorded once by stream capture and replayed per
// iteration.  What changes between iterations is the seed: the graph starts
// with a one-block kernel that writes the stream keys of the new seed into a
// device table, and every kernel of the pass reads its key from that table.
// The graph owns its scratch (allocated once, at capture) and its own
// work-counter blocks, which the trace kernels reset as they finish, so a
// replay needs no memset node.  One graph replays on one stream at a time.
#include "rb_device.h"

namespace rb {

constexpr uint32_t kKeyTags = 16;

__global__ void k_set_keys(Rng* keys, uint64_t seed) {
    const uint32_t t = threadIdx.x;
    if (t < kKeyTags) keys[t] = Rng::make(seed, t);
}

}  // namespace rb

using namespace rb;

struct rb_render_graph_s {
    hipGraph_t body = nullptr, top = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t keys_node = nullptr;
    Rng* keys = nullptr;
    uint32_t* counters = nullptr;
    std::vector<void*> allocs;
    hipKernelNodeParams kp{};
};

static void graph_free(rb_render_graph_s* g) {
    if (!g) return;
    (void)hipDeviceSynchronize();  // no replay may still use the scratch
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->top) (void)hipGraphDestroy(g->top);
    if (g->body) (void)hipGraphDestroy(g->body);
    for (void* p : g->allocs) (void)hipFree(p);
    if (g->keys) (void)hipFree(g->keys);
    if (g->counters) (void)hipFree(g->counters);
    delete g;
}

extern "C" {

rb_status rb_render_graph_create(uint32_t kind, rb_bvh_t bvh, const rb_camera* cam, uint32_t x0, uint32_t y0,
                                 uint32_t w, uint32_t h, const rb_render_params* p, float* out_rgb,
                                 rb_render_graph_t* out) {
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    if (!out || !p) return fail(RB_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (kind != RB_REQ_AO && kind != RB_REQ_PT && kind != RB_REQ_BDPT && kind != RB_REQ_RAYCAST)
        return fail(RB_ERR_INVALID, "render graphs take RB_REQ_AO, RB_REQ_PT, RB_REQ_BDPT or RB_REQ_RAYCAST");
    if (p->ray_counter || p->path_counter || p->light_path_counter)
        return fail(RB_ERR_INVALID, "render graphs take no ray or path counters");
    rb_render_graph_s* G = new rb_render_graph_s();
    const uint32_t pairs = 2 * (p->max_depth + 1) + p->light_max_depth + 8;
    hipError_t e;
    if ((e = hipMalloc((void**)&G->keys, kKeyTags * sizeof(Rng))) != hipSuccess ||
        (e = hipMalloc((void**)&G->counters, pairs * kCounterWords * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMemset(G->counters, 0, pairs * kCounterWords * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMemset(G->keys, 0, kKeyTags * sizeof(Rng))) != hipSuccess) {
        graph_free(G);
        return hip_fail(e, "render graph setup");
    }
    hipStream_t cs = nullptr;
    if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess) {
        graph_free(G);
        return hip_fail(e, "render graph stream");
    }
    CaptureCtx ctx;
    ctx.keys = G->keys;
    ctx.counters = G->counters;
    ctx.n_pairs = pairs;
    if ((e = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed)) != hipSuccess) {
        (void)hipStreamDestroy(cs);
        graph_free(G);
        return hip_fail(e, "hipStreamBeginCapture");
    }
    t_capture = &ctx;
    switch (kind) {
        case RB_REQ_AO: st = render_ao(bvh, cam, x0, y0, w, h, p, out_rgb, cs); break;
        case RB_REQ_PT: st = render_pt(bvh, cam, x0, y0, w, h, p, out_rgb, cs); break;
        case RB_REQ_RAYCAST: st = render_raycast(bvh, cam, x0, y0, w, h, p, out_rgb, cs); break;
        default: st = render_bdpt(bvh, cam, x0, y0, w, h, p, out_rgb, cs); break;
    }
    t_capture = nullptr;
    hipGraph_t body = nullptr;
    e = hipStreamEndCapture(cs, &body);
    (void)hipStreamDestroy(cs);
    G->allocs = std::move(ctx.allocs);
    if (st != RB_OK || e != hipSuccess || !body) {
        if (body) (void)hipGraphDestroy(body);
        graph_free(G);
        return st != RB_OK ? st : hip_fail(e, "hipStreamEndCapture");
    }
    G->body = body;
    static uint64_t seed0 = 0;
    void* args[2] = {&G->keys, &seed0};
    G->kp.func = reinterpret_cast<void*>(k_set_keys);
    G->kp.gridDim = dim3(1);
    G->kp.blockDim = dim3(64);
    G->kp.sharedMemBytes = 0;
    G->kp.kernelParams = args;
    G->kp.extra = nullptr;
    hipGraphNode_t child = nullptr;
    if ((e = hipGraphCreate(&G->top, 0)) != hipSuccess ||
        (e = hipGraphAddKernelNode(&G->keys_node, G->top, nullptr, 0, &G->kp)) != hipSuccess ||
        (e = hipGraphAddChildGraphNode(&child, G->top, &G->keys_node, 1, G->body)) != hipSuccess ||
        (e = hipGraphInstantiate(&G->exec, G->top, nullptr, nullptr, 0)) != hipSuccess) {
        graph_free(G);
        return hip_fail(e, "render graph instantiate");
    }
    G->kp.kernelParams = nullptr;
    *out = G;
    return RB_OK;
}

rb_status rb_render_graph_launch(rb_render_graph_t g, uint64_t seed, void* stream) {
    if (!g) return fail(RB_ERR_INVALID, "NULL graph");
    rb_status st = ensure_device();
    if (st != RB_OK) return st;
    hipKernelNodeParams kp = g->kp;
    void* args[2] = {&g->keys, &seed};
    kp.kernelParams = args;
    hipError_t e;
    if ((e = hipGraphExecKernelNodeSetParams(g->exec, g->keys_node, &kp)) != hipSuccess)
        return hip_fail(e, "render graph seed update");
    if ((e = hipGraphLaunch(g->exec, stream ? (hipStream_t)stream : dev().stream)) != hipSuccess)
        return hip_fail(e, "hipGraphLaunch");
    return RB_OK;
}

rb_status rb_render_graph_destroy(rb_render_graph_t g) {
    graph_free(g);
    return RB_OK;
}

}  // extern "C"
