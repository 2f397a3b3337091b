#!/usr/bin/env python3
"""Benchmark: incoherent closest-hit Mrays/s on the 1M-triangle trace scene.

Workload (one "step" = one rb_trace launch over one batch):
  scene   1,000,000 triangles = 750 icospheres (subdivision 3) + 20,000 thin
          1:50 quads in [-10,10]^3, seed 0, SBVH with the reference defaults
  rays    2,097,152 incoherent diffuse rays: origins on the hits of a
          2048x1024 pinhole view, cosine-distributed about the geometric
          normal, tmin = 1e-3 t  (the headline set); the 2,097,152 coherent
          primaries and the mixed 4,194,304 set are reported beside it
  output  (triangle id, t) per ray, closest hit
Inputs are resident in HBM before timing starts; HIP events on the trace
stream bracket exactly `steps` launches.  One JSON line goes to stdout.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "incoherent closest-hit Mrays/s @1M tris; batched vs plain; p50/p99 latency"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def check_tree_fixture(rb, scene, bvh):
    """The 1M-triangle tree this run built (on every CPU) must be the bytes of
    the single-threaded oracle's tree recorded in tests/golden/trace_1m_tree.json:
    the algorithmic bytes per ray are counted on it.  Exits on a mismatch."""
    import hashlib

    import numpy as np
    from renderbox2_amd import _lib as L
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "trace_1m_tree.json")))
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    v, idx, _, _ = scene.geometry()
    n, w, x = bvh.export(L.RB_LAYOUT_REF_COMPACT2)
    got = (sha(v), sha(idx), sha(n), sha(w), sha(x))
    want = (g["sha_verts"], g["sha_idx"], g["sha_nodes"], g["sha_woop"], g["sha_index"])
    if got != want:
        raise SystemExit(f"bench: the 1M-triangle scene or its SBVH differs from the oracle's fixture "
                         f"(scene {got[:2] == want[:2]}, tree {got[2:] == want[2:]}); results would not be comparable")
    return "sha256 " + g["sha_nodes"][:16]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_launches(rb, tracer, d_rays, n, d_hits, any_hit, steps, warmup, stream):
    for _ in range(warmup):
        tracer.trace_device(d_rays, n, d_hits, any_hit=any_hit, stream=stream)
    stream.synchronize()
    e0, e1 = rb.Event(), rb.Event()
    e0.record(stream)
    for _ in range(steps):
        tracer.trace_device(d_rays, n, d_hits, any_hit=any_hit, stream=stream)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1) / steps


def read_profile_value(path, key):
    """A `key = value` figure from a committed rocprofv3 --pmc summary, if present."""
    try:
        d = {}
        for line in open(path):
            if "=" in line and not line.startswith("#"):
                k, v = line.split("=", 1)
                d[k.strip()] = v.strip()
        return float(d[key])
    except (OSError, KeyError, ValueError):
        return None


def read_traffic(path):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    return read_profile_value(path, "traffic_bytes_per_launch")


def bench_render(rb, kind, scene, bvh_max_leaf, cam, frames, built=None, **params):
    """Time whole frames (raygen + integrator + film) with HIP events; rays
    traced per frame come from the device ray counter.  built: (tracer,
    build seconds) of `scene` when already uploaded."""
    import ctypes as C

    import numpy as np
    if built is None:
        t0 = time.time()
        bvh = rb.scenes.build(scene, max_leaf=bvh_max_leaf)
        built = (rb.CudaTracer(bvh, scene), time.time() - t0)
    tracer, build_s = built
    stream = rb.Stream()
    counter = rb.DeviceBuffer(8)
    counter.zero()
    r = rb.SamplerRenderer(tracer, cam, kind, **params)
    r.params.ray_counter = C.c_void_p(counter.ptr)
    # live camera-path vertices per bounce and light sub-paths per depth
    depth = r.params.max_depth + 1
    paths = rb.DeviceBuffer(8 * depth)
    paths.zero()
    lpaths = rb.DeviceBuffer(8 * max(r.params.light_max_depth, 1))
    lpaths.zero()
    if kind in ("pt", "bdpt"):
        r.params.path_counter = C.c_void_p(paths.ptr)
    if kind == "bdpt":
        r.params.light_path_counter = C.c_void_p(lpaths.ptr)
    out = rb.DeviceBuffer(cam.width * cam.height * 12)
    r.render_device(out, stream=stream.handle)
    stream.synchronize()
    rays = int(counter.download(np.uint64, 1)[0])
    r.params.ray_counter = r.params.path_counter = r.params.light_path_counter = None
    # one event pair per frame: the median frame, and the mean over all
    ev = [rb.Event() for _ in range(frames + 1)]
    ev[0].record(stream)
    for k in range(frames):
        r.render_device(out, stream=stream.handle)
        ev[k + 1].record(stream)
    ev[-1].synchronize()
    per = [ev[k].elapsed_ms(ev[k + 1]) for k in range(frames)]
    ms = float(np.median(per))
    img = out.download(np.float32).reshape(cam.height, cam.width, 3)
    res = {"frame_ms": round(ms, 3), "frame_ms_mean": round(float(np.mean(per)), 3), "frames": frames,
           "rays_per_frame": rays, "mrays": round(rays / (ms * 1e-3) / 1e6, 1),
           "tris": scene.get_num_triangles(), "bvh_build_s": round(build_s, 1), "mean": round(float(img.mean()), 5)}
    if kind in ("pt", "bdpt"):
        res["paths_per_bounce"] = [int(x) for x in paths.download(np.uint64, depth)]
    if kind == "bdpt":
        res["light_paths_per_depth"] = [int(x) for x in lpaths.download(np.uint64, r.params.light_max_depth)]
    return res


def bench_graphs(rb, built, iters=50):
    """Progressive rendering of one tile, `iters` iterations with a new seed
    each: eager (every kernel launched from the host per iteration) vs the
    pass captured once into a HIP graph and replayed (rb_render_graph_*).
    Small tiles are where launch overhead shows."""
    import numpy as np
    tracer = built[0]
    cam = rb.scenes.pt_camera(1024, 1024)
    stream = rb.Stream()
    out = {}
    cases = [("ao_64x64", "ao", (480, 480, 64, 64), dict(spp=1, ao_samples=8)),
             ("pt_128x128", "pt", (448, 448, 128, 128), dict(spp=1, max_depth=8)),
             ("bdpt_128x128", "bdpt", (448, 448, 128, 128),
              dict(spp=1, max_depth=6, light_max_depth=6, light_paths=16384, connections=5))]
    for name, kind, tile, p in cases:
        d = rb.DeviceBuffer(tile[2] * tile[3] * 12)
        r = rb.SamplerRenderer(tracer, cam, kind, **p)
        g = rb.RenderGraph(tracer, cam, kind, d, tile, **p)

        def eager():
            for i in range(iters):
                r.params.seed = i
                r.render_device(d, tile, stream.handle)

        def graph():
            for i in range(iters):
                g.launch(i, stream.handle)

        res = {}
        for tag, fn in (("eager", eager), ("graph", graph)):
            fn()
            stream.synchronize()
            e0, e1 = rb.Event(), rb.Event()
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            res[tag] = e0.elapsed_ms(e1) / iters
        img_g = d.download(np.float32)
        r.params.seed = iters - 1
        r.render_device(d, tile, stream.handle)
        stream.synchronize()
        same = bool(np.array_equal(img_g.view(np.uint32), d.download(np.float32).view(np.uint32)))
        g.close()
        out[name] = {"eager_ms": round(res["eager"], 4), "graph_ms": round(res["graph"], 4),
                     "speedup": round(res["eager"] / res["graph"], 3), "match": same}
    return out


BATCH_SCENES = [10_000, 50_000, 200_000, 500_000, 1_000_000, 2_000_000, 3_000_000, 4_000_000]


def bench_batched(rb, reps=5, seed=0):
    """Config 4: 8 resident scenes (trace generator, seeds 0-7), 64 closest-hit
    batches of log-uniform size in [4096, 2 Mi] rays, half primaries / half
    diffuse rays of scene (batch % 8).  plain = one rb_trace + stream sync per
    batch; batched = one rb_trace_batch call + one sync."""
    import numpy as np

    from renderbox2_amd import _lib as L
    t0 = time.time()
    stream = rb.Stream()
    tracers, pools = [], []
    cam = rb.scenes.trace_camera(1024, 1024)
    npool = cam.get_total_pixels()
    for k, T in enumerate(BATCH_SCENES):
        s = rb.scenes.trace_scaled(T, seed=k)
        tr = rb.CudaTracer(rb.scenes.build(s), s)
        prim, diff, h = rb.DeviceBuffer(npool * 32), rb.DeviceBuffer(npool * 32), rb.DeviceBuffer(npool * 8)
        cam.generate_rays_device(prim, seed=k, stream=stream)
        tr.trace_device(prim, npool, h, stream=stream)
        rb.raygen_diffuse_device(tr, prim, h, npool, npool, 100 + k, diff, stream=stream)
        stream.synchronize()
        tracers.append(tr)
        pools.append((prim, diff))
        log(f"batched: scene {k} ({T} tris) ready, {time.time() - t0:.1f}s")
    build_s = time.time() - t0
    rng = np.random.default_rng(seed)
    sizes = (4096 * 512.0 ** rng.uniform(0, 1, 64)).astype(np.int64) & ~1
    batches = []
    for i, sz in enumerate(sizes):
        k = i % 8
        half = int(sz) // 2
        d = rb.DeviceBuffer(int(sz) * 32)
        o1, o2 = (int(x) for x in rng.integers(0, npool - half + 1, 2))
        L.check(L.lib().rb_memcpy_dtod(d.ptr, pools[k][0].ptr + o1 * 32, half * 32, stream.handle))
        L.check(L.lib().rb_memcpy_dtod(d.ptr + half * 32, pools[k][1].ptr + o2 * 32, half * 32, stream.handle))
        batches.append((tracers[k], d, int(sz), rb.DeviceBuffer(int(sz) * 8), False))
    stream.synchronize()
    total = int(sizes.sum())

    def plain():
        for tr, d, n, h, _ in batches:
            tr.trace_device(d, n, h, stream=stream)
            stream.synchronize()

    def batched():
        rb.trace_batch(batches, stream=stream)
        stream.synchronize()

    plain()
    ref = [h.download(np.uint32, 2 * n) for _, _, n, h, _ in batches]
    batched()
    match = all(np.array_equal(r, h.download(np.uint32, 2 * n)) for r, (_, _, n, h, _) in zip(ref, batches))
    tp, tb = [], []
    for _ in range(reps):
        t1 = time.perf_counter()
        plain()
        tp.append(time.perf_counter() - t1)
        t1 = time.perf_counter()
        batched()
        tb.append(time.perf_counter() - t1)
    tp, tb = sorted(tp)[reps // 2], sorted(tb)[reps // 2]
    return {"plain_mrays": round(total / tp / 1e6, 1), "batched_mrays": round(total / tb / 1e6, 1),
            "speedup": round(tp / tb, 3), "plain_ms": round(tp * 1e3, 2), "batched_ms": round(tb * 1e3, 2),
            "rays": total, "batches": len(batches), "scene_tris": BATCH_SCENES, "match": bool(match),
            "setup_s": round(build_s, 1)}


SERVED_SCENES = [100_000, 500_000, 1_000_000, 2_000_000]


def served_mix(nreq, seed, pt_cam, bd_cam):
    """Config 5 request list: 40% AO tiles (side 64-512, 8 AO rays), 40% PT
    tiles (side 128-1024, 4-16 spp, depth 8), 20% BDPT frames (512x512,
    4 spp, light-vertex cache of 2^20 vertices, depth 6), scenes uniform."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nreq):
        u, sc = rng.uniform(), int(rng.integers(0, len(SERVED_SCENES)))
        rs = 10_000 + i
        if u < 0.4:
            side = 8 * int(rng.integers(8, 65))
            x0, y0 = (int(v) for v in rng.integers(0, pt_cam.width - side + 1, 2))
            out.append(("ao", sc, pt_cam, (x0, y0, side, side), dict(spp=1, seed=rs, ao_samples=8, ao_radius=1.0)))
        elif u < 0.8:
            side = 8 * int(rng.integers(16, 129))
            x0, y0 = (int(v) for v in rng.integers(0, pt_cam.width - side + 1, 2))
            out.append(("pt", sc, pt_cam, (x0, y0, side, side),
                        dict(spp=int(rng.integers(4, 17)), seed=rs, max_depth=8, rr_start=3, rr_prob=0.5)))
        else:
            out.append(("bdpt", sc, bd_cam, None, dict(spp=4, seed=rs, max_depth=6, rr_start=3, rr_prob=0.5,
                                                       light_max_depth=6, lvc_capacity=1 << 20, light_paths=0,
                                                       connections=5)))
    return out


def _digest(img):
    try:
        import xxhash
        return xxhash.xxh3_64_intdigest(img.tobytes())
    except ImportError:
        import hashlib
        return hashlib.blake2b(img.tobytes(), digest_size=8).hexdigest()


def bench_served(rb, nreq=2000, load=0.8, max_inflight=8, policy="deadline", seed=0):
    """Config 5.  The same seeded request list three ways:
    plain  : one request at a time on one stream, host sync + result copy
             after each (the reference's synchronous call pattern);
    burst  : every request submitted to the server at once -> capacity;
    poisson: arrivals at `load` x capacity -> p50/p99 latency (submit ->
             result on the host) and aggregate Mrays/s.
    The server runs burst and poisson in both modes: each request on its own
    stream (the top-level figures) and in lockstep (one batched trace launch
    of every in-flight request's rays per round; "lockstep"), the poisson
    arrivals of both at `load` x the stream mode's capacity.
    Rays per request come from the device ray counter; every served image is
    checked against the plain one by digest."""
    import ctypes as C
    import queue
    import threading

    import numpy as np
    t0 = time.time()
    tracers, tris = [], []
    for k, T in enumerate(SERVED_SCENES):
        s = rb.scenes.pt_scaled(T, seed=k)
        tracers.append(rb.CudaTracer(rb.scenes.build(s), s))
        tris.append(s.get_num_triangles())
        log(f"served: scene {k} ({s.get_num_triangles()} tris) ready, {time.time() - t0:.1f}s")
    setup_s = time.time() - t0
    pt_cam, bd_cam = rb.scenes.pt_camera(1024, 1024), rb.scenes.pt_camera(512, 512)
    reqs = served_mix(nreq, seed, pt_cam, bd_cam)

    # plain
    stream = rb.Stream()
    counter = rb.DeviceBuffer(8)
    dout = rb.DeviceBuffer(1024 * 1024 * 12)
    ref, rays = [], []
    t_plain = 0.0
    tl = time.time()
    for i, (kind, sc, cam, tile, p) in enumerate(reqs):
        r = rb.SamplerRenderer(tracers[sc], cam, kind, **p)
        r.params.ray_counter = C.c_void_p(counter.ptr)
        counter.zero(stream.handle)
        stream.synchronize()
        t1 = time.perf_counter()
        w, h = r.render_device(dout, tile, stream.handle)
        img = dout.download(np.float32, w * h * 3, stream.handle)
        stream.synchronize()
        t_plain += time.perf_counter() - t1
        ref.append(_digest(img))
        rays.append(int(counter.download(np.uint64, 1)[0]))
        if time.time() - tl > 10:
            log(f"served/plain: {i + 1}/{nreq} requests, {t_plain:.1f}s of work")
            tl = time.time()
    total_rays = sum(rays)
    kinds = [q[0] for q in reqs]

    def run(arrivals, mode="streams"):
        srv = rb.RenderServer(max_inflight=max_inflight, policy=policy, mode=mode)
        sids = [srv.add_scene(t) for t in tracers]
        rq = queue.Queue()

        def submitter():
            ts = time.perf_counter()
            for i, (kind, sc, cam, tile, p) in enumerate(reqs):
                if arrivals is not None:
                    dt = arrivals[i] - (time.perf_counter() - ts)
                    if dt > 0:
                        time.sleep(dt)
                rq.put(srv.submit_render(kind, sids[sc], cam, tile, **p))

        th = threading.Thread(target=submitter)
        th.start()
        ok, lat, got_rays = True, [], 0
        tl = time.time()
        for i in range(nreq):
            st, img = srv.wait(rq.get())
            ok &= st.result == 0 and _digest(img) == ref[i]
            lat.append(st.latency_ms)
            got_rays += st.rays
            if time.time() - tl > 10:
                log(f"served/{mode}/{'poisson' if arrivals is not None else 'burst'}: {i + 1}/{nreq} done")
                tl = time.time()
        th.join()
        stt = srv.stats()
        srv.close()
        return ok and got_rays == total_rays, np.array(lat), stt

    ok_b, lat_b, st_b = run(None)
    span_b = st_b["first_submit_to_last_done_s"]
    cap = nreq / span_b
    arr = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / (load * cap), nreq))
    ok_p, lat_p, st_p = run(arr)
    span_p = st_p["first_submit_to_last_done_s"]
    ok_lb, lat_lb, st_lb = run(None, "lockstep")
    span_lb = st_lb["first_submit_to_last_done_s"]
    ok_lp, lat_lp, st_lp = run(arr, "lockstep")  # the same arrivals
    span_lp = st_lp["first_submit_to_last_done_s"]
    lockstep = {"served_capacity_mrays": round(total_rays / span_lb / 1e6, 1), "burst_s": round(span_lb, 2),
                "speedup": round(t_plain / span_lb, 3), "poisson_rate_per_s": round(load * cap, 2),
                "poisson_mrays": round(total_rays / span_lp / 1e6, 1), "poisson_s": round(span_lp, 2),
                "p50_ms": round(st_lp["p50_ms"], 2), "p99_ms": round(st_lp["p99_ms"], 2),
                "burst_p50_ms": round(st_lb["p50_ms"], 2), "burst_p99_ms": round(st_lb["p99_ms"], 2),
                "arena_misses": st_lp["arena_misses"], "match": bool(ok_lb and ok_lp)}

    def pct(a, q):
        return round(float(np.percentile(a, q)), 2) if len(a) else None

    per_kind = {}
    for k in ("ao", "pt", "bdpt"):
        sel = np.array([x == k for x in kinds])
        per_kind[k] = {"n": int(sel.sum()), "p50_ms": pct(lat_p[sel], 50), "p99_ms": pct(lat_p[sel], 99),
                       "mrays_each": round(sum(r for r, x in zip(rays, kinds) if x == k) / max(1, sel.sum()) / 1e6, 2)}
    return {"requests": nreq, "scene_tris": tris, "rays": total_rays,
            "plain_mrays": round(total_rays / t_plain / 1e6, 1), "plain_s": round(t_plain, 2),
            "served_capacity_mrays": round(total_rays / span_b / 1e6, 1), "burst_s": round(span_b, 2),
            "speedup": round(t_plain / span_b, 3),
            "poisson_load": load, "poisson_rate_per_s": round(load * cap, 2),
            "poisson_mrays": round(total_rays / span_p / 1e6, 1), "poisson_s": round(span_p, 2),
            "p50_ms": round(st_p["p50_ms"], 2), "p99_ms": round(st_p["p99_ms"], 2),
            "burst_p50_ms": round(st_b["p50_ms"], 2), "burst_p99_ms": round(st_b["p99_ms"], 2),
            "per_kind": per_kind, "max_inflight": max_inflight, "policy": policy, "deferrals": st_p["deferrals"],
            "arena_misses": st_p["arena_misses"], "peak_arena_mb": round(st_p["peak_arena_bytes"] / 2**20, 1),
            "match": bool(ok_b and ok_p), "setup_s": round(setup_s, 1), "lockstep": lockstep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline budget (0 = skip)")
    ap.add_argument("--extras", type=int, default=1, help="also time coherent / any-hit / mixed sets")
    ap.add_argument("--autotune", type=int, default=0,
                    help="1: pick the traversal config on this batch first (default 0: the library's default "
                         "config, which the autotuner also settles on; its timing noise picked configs up to "
                         "1%% slower)")
    ap.add_argument("--config", default="",
                    help="block,lds_stack,chunk,fetch_threshold,waves_per_cu[,speculative[,leaf_gate[,lds_nodes[,wide]]]]")
    ap.add_argument("--speculative", type=int, default=0,
                    help="1: the reference's speculative leaf order (wave-dependent results in rare cases)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="launch one copy kernel of this size first (traffic counter calibration)")
    ap.add_argument("--render", type=int, default=1, help="also time the AO and path-traced frames")
    ap.add_argument("--graphs", type=int, default=1, help="with --render: also time hipGraph replay of tiles")
    ap.add_argument("--batched", type=int, default=1, help="also time batched vs plain (8 scenes, 64 batches)")
    ap.add_argument("--served", type=int, default=2000, help="served-mix requests (0 = skip)")
    ap.add_argument("--served-inflight", type=int, default=8)
    ap.add_argument("--served-policy", default="deadline", choices=["deadline", "fifo"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(rays shard by replicating the scene and splitting each batch; no figure at several GPUs is "
              "produced here)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import renderbox2_amd as rb
    from renderbox2_amd import _lib as L

    rb.device_init(0)
    import shutil
    import tempfile
    scene = rb.scenes.trace_1m(seed=0)
    cache_dir = tempfile.mkdtemp(prefix="rbvh_")
    try:
        t0 = time.time()
        bvh, _ = rb.BVH.build_cached(scene, cache_dir)  # miss: build + write the binary cache
        build_s = time.time() - t0
        t0 = time.time()
        bvh, hit = rb.BVH.build_cached(scene, cache_dir)  # hit: load and check the key
        load_s = time.time() - t0
        cache_mb = sum(os.path.getsize(os.path.join(cache_dir, f)) for f in os.listdir(cache_dir)) / 1e6
    finally:
        shutil.rmtree(cache_dir, ignore_errors=True)
    st = bvh.stats()
    log(f"scene {scene.get_num_triangles()} tris, SBVH {st['n_inner']} inner / {st['n_tri_refs']} refs "
        f"in {build_s:.1f}s; binary cache {cache_mb:.1f} MB loads in {load_s:.3f}s (hit={hit})")
    tree_check = check_tree_fixture(rb, scene, bvh)
    log(f"1M tree: bytes equal the single-threaded oracle's ({tree_check})")
    tracer = rb.CudaTracer(bvh, scene)
    cam = rb.scenes.trace_camera(2048, 1024)
    n = cam.get_total_pixels()
    stream = rb.Stream()
    d_prim = rb.DeviceBuffer(n * 32)
    d_diff = rb.DeviceBuffer(n * 32)
    d_mixed = rb.DeviceBuffer(2 * n * 32)
    d_hits = rb.DeviceBuffer(2 * n * 8)
    cam.generate_rays_device(d_prim, spp=1, seed=0, jitter=True, stream=stream)
    tracer.trace_device(d_prim, n, d_hits, stream=stream)
    rb.raygen_diffuse_device(tracer, d_prim, d_hits, n, n, 7, d_diff, stream=stream)
    L.check(L.lib().rb_memcpy_dtod(d_mixed.ptr, d_prim.ptr, n * 32, stream.handle))
    L.check(L.lib().rb_memcpy_dtod(d_mixed.ptr + n * 32, d_diff.ptr, n * 32, stream.handle))
    stream.synchronize()
    if args.copy_probe_mib:
        nb = args.copy_probe_mib << 20
        src, dst = rb.DeviceBuffer(nb), rb.DeviceBuffer(nb)
        src.zero(stream.handle)
        rb.copy_kernel(dst, src, nb, stream)
        stream.synchronize()
        del src, dst
    tuned_ms = None
    rb.set_trace_config(speculative=args.speculative)
    if args.config:
        keys = ["block", "lds_stack", "chunk", "fetch_threshold", "waves_per_cu", "speculative", "leaf_gate",
                "lds_nodes", "wide"]
        rb.set_trace_config(**dict(zip(keys, (int(x) for x in args.config.split(",")))))
    elif args.autotune:
        _, tuned_ms = rb.autotune(tracer, d_diff, n, d_hits, reps=5, stream=stream)
    cfg = rb.get_trace_config()
    log("trace config", cfg, "autotuned %.3f ms" % tuned_ms if tuned_ms else "")

    ms = time_launches(rb, tracer, d_diff, n, d_hits, False, args.steps, args.warmup, stream)
    value = n / (ms * 1e-3) / 1e6
    log(f"incoherent closest: {ms:.3f} ms/launch, {value:.1f} Mrays/s")
    hits = d_hits.download(rb.HIT_DTYPE, n)
    hit_frac = float((hits["id"] >= 0).mean())

    extra = {}
    if args.extras:
        for name, buf, cnt, anyh in (("coherent_closest", d_prim, n, False), ("incoherent_any", d_diff, n, True),
                                     ("mixed_closest", d_mixed, 2 * n, False)):
            m = time_launches(rb, tracer, buf, cnt, d_hits, anyh, max(3, args.steps // 2), 1, stream)
            extra[name + "_mrays"] = round(cnt / (m * 1e-3) / 1e6, 1)
            log(f"{name}: {m:.3f} ms/launch, {extra[name + '_mrays']} Mrays/s")
        # one GPU's share of the mixed 4 Mi set split over G GPUs (every G-th
        # ray, the interleaved split), traced at its real size on this GPU
        mixed = d_mixed.download(np.float32, 2 * n * 8).reshape(2 * n, 8)
        share = {}
        for G in (2, 4, 8):
            part = np.ascontiguousarray(mixed[::G])
            d_part = rb.DeviceBuffer.from_array(part)
            m = time_launches(rb, tracer, d_part, len(part), d_hits, False, max(3, args.steps // 2), 1, stream)
            share[str(G)] = {"rays": len(part), "ms": round(m, 4), "mrays": round(len(part) / (m * 1e-3) / 1e6, 1)}
            del d_part
        extra["mixed_share_per_gpu"] = share
        log(f"mixed set, one GPU's share: {share}")
        del mixed

    # device-side traversal counters of the binary walk (the reference's
    # visit order, priced with the reference's formula), all rays
    ids, _, gcnt = tracer.trace_counts(d_diff.download(np.float32, n * 8).reshape(n, 8))
    gpu_bpr = float((40 + 64 * gcnt[:, 0] + 48 * gcnt[:, 1] + 16 * gcnt[:, 2] + 4 * (ids >= 0)).mean())
    # bytes the timed kernel itself reads and writes per ray: the ray (32),
    # the hit (8), each visited record (64, binary or wide), each tested
    # triangle (48) and the hit's index entry (4)
    wide = bool(cfg["wide"]) and not cfg["speculative"]
    if wide:
        wids, _, wcnt = tracer.trace_counts(d_diff.download(np.float32, n * 8).reshape(n, 8), wide=True)
        kernel_bpr = float((40 + 64 * wcnt[:, 0] + 48 * wcnt[:, 1] + 4 * (wids >= 0)).mean())
        wide_records = float(wcnt[:, 0].mean())
        del wids, wcnt
    else:
        kernel_bpr = float((40 + 64 * gcnt[:, 0] + 48 * gcnt[:, 1] + 4 * (ids >= 0)).mean())
        wide_records = None

    # CPU baseline (the oracle's restatement of the reference traversal) and
    # the algorithmic bytes per ray it counts on the reference visit order.
    cpu = None
    bytes_per_ray = None
    if args.cpu_seconds > 0:
        from oracle import oracle
        # every CPU this process may run on, capped by the host's own thread
        # setting where it has one (the GPU host gives a GPU's job 16 CPUs)
        affinity = len(os.sched_getaffinity(0))
        omp = os.environ.get("OMP_NUM_THREADS", "")
        cores = min(affinity, int(omp)) if omp.isdigit() and int(omp) > 0 else affinity
        rays = d_diff.download(np.float32, n * 8).reshape(n, 8)
        cn, cw, cx = bvh.export(L.RB_LAYOUT_REF_COMPACT2)
        cn, cw = cn.reshape(-1, 4), cw.reshape(-1, 4)
        probe = rays[: 1 << 15]
        t1 = time.time()
        oracle.trace_compact2(cn, cw, cx, probe, nthreads=cores)
        per_ray = (time.time() - t1) / len(probe)
        # 1 warm-up + 5 timed runs within the budget
        size = int(min(n, max(1 << 16, args.cpu_seconds / 6.0 / max(per_ray, 1e-9))))
        sample = rays[:size]
        times = []
        for rep in range(6):
            t1 = time.time()
            oid, ot, ocnt = oracle.trace_compact2(cn, cw, cx, sample, counts=True, nthreads=cores)
            if rep:
                times.append(time.time() - t1)
        tcpu = sorted(times)[2]
        bytes_per_ray = float(oracle.bytes_per_ray(ocnt, oid >= 0).mean())
        counts_equal = bool(np.array_equal(ocnt, gcnt[:size]))
        same = bool(np.array_equal(oid, hits["id"][:size]) and
                    np.array_equal(ot.view(np.uint32), hits["t"][:size].view(np.uint32)))
        cpu = {"value": round(size / tcpu / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
               "sample": f"first {size} of the {n} incoherent rays, closest hit, 1 warm-up + median of 5, "
                         f"OpenMP {cores} threads ({affinity} CPUs in the affinity mask, OMP_NUM_THREADS="
                         f"{omp or 'unset'}), {cpu_model()}",
               "cpu_model": cpu_model(), "affinity_cpus": affinity,
               "matches_gpu": same, "visit_counts_equal": counts_equal}
        log(f"cpu baseline: {cpu['value']} Mrays/s on {cores} threads ({affinity} CPUs), B/ray {bytes_per_ray:.0f}, "
            f"match={same}")
    if bytes_per_ray is None:
        bytes_per_ray = gpu_bpr

    if args.render or args.batched or args.served:
        del tracer, d_prim, d_diff, d_mixed, d_hits
    if args.batched:
        bt = bench_batched(rb)
        log("batched vs plain (config 4)", bt)
        extra["batched"] = bt
    if args.render:
        ao = bench_render(rb, "ao", rb.scenes.ao_room(seed=0), 4, rb.scenes.ao_camera(512, 512), 10,
                          spp=1, seed=0, ao_samples=8, ao_radius=1.0)
        log("ao frame (config 1: 78,848 tris, 512x512, 1 spp, 8 AO rays)", ao)
        extra["ao_512"] = ao
        room = rb.scenes.pt_room(seed=0)
        t0 = time.time()
        built = (rb.CudaTracer(rb.scenes.build(room), room), time.time() - t0)
        pt = bench_render(rb, "pt", room, None, rb.scenes.pt_camera(1024, 1024), 7, built,
                          spp=8, seed=0, max_depth=8, rr_start=3, rr_prob=0.5)
        log("pt frame (config 3: 2,000,136 tris, 1024x1024, 8 spp, depth 8)", pt)
        extra["pt_1024_8spp"] = pt
        bd = bench_render(rb, "bdpt", room, None, rb.scenes.pt_camera(512, 512), 9, built,
                          spp=4, seed=0, max_depth=6, rr_start=3, rr_prob=0.5, light_max_depth=6,
                          lvc_capacity=1 << 20, light_paths=0, connections=5)
        log("bdpt frame (config 5 BDPT: 2,000,136 tris, 512x512, 4 spp, 2^20-vertex cache, depth 6)", bd)
        extra["bdpt_512_4spp"] = bd
        if args.graphs:
            gr = bench_graphs(rb, built)
            log("progressive tiles, eager vs hipGraph replay (2,000,136 tris)", gr)
            extra["graph_replay"] = gr
        del built
    if args.served:
        sv = bench_served(rb, args.served, max_inflight=args.served_inflight, policy=args.served_policy)
        log("served mix (config 5)", sv)
        extra["served"] = sv

    # achieved: the reference traversal's algorithmic bytes over the timed
    # kernel's time; kernel_achieved: the kernel's own reads and writes
    achieved = bytes_per_ray * n / (ms * 1e-3) / 1e9
    kernel_achieved = kernel_bpr * n / (ms * 1e-3) / 1e9
    traffic = read_traffic(os.path.join(ROOT, "profiles", "trace_incoherent_1m_pmc.txt"))
    l2_hit = read_profile_value(os.path.join(ROOT, "profiles", "trace_incoherent_1m_l2.txt"), "l2_hit_rate")
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "trace_1m incoherent closest-hit: 1,000,000 tris (750 icospheres s3 + "
                               "20,000 thin quads), 2,097,152 diffuse rays from 2048x1024 primary hits",
                   "tris": scene.get_num_triangles(), "rays": n, "hit_fraction": round(hit_frac, 4),
                   "bvh_inner_nodes": st["n_inner"], "bvh_tri_refs": st["n_tri_refs"],
                   "bvh_build_s": round(build_s, 1), "bvh_cache_load_s": round(load_s, 3),
                   "bvh_cache_mb": round(cache_mb, 1), "trace_config": cfg,
                   "bvh_equals_oracle_tree": tree_check},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_ray": round(bytes_per_ray, 1), "gpu_visit_bytes_per_ray": round(gpu_bpr, 1),
                     "kernel": "k_trace_wide" if wide else "k_trace",
                     "kernel_bytes_per_ray": round(kernel_bpr, 1), "kernel_achieved": round(kernel_achieved, 1),
                     "kernel_frac": round(kernel_achieved / HBM_PEAK_GBS, 4),
                     "wide_records_per_ray": None if wide_records is None else round(wide_records, 2),
                     "binary_records_per_ray": round(float(gcnt[:, 0].mean()), 2),
                     "tris_per_ray": round(float(gcnt[:, 1].mean()), 2),
                     "l2_hit_rate": l2_hit},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
