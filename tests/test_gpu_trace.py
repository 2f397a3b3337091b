"""GPU parity of the trace path: librb.so's HIP kernels against the oracle.

Closest hit: triangle id and t bit-identical to the oracle's restatement of
the reference traversal on the same SBVH (ties to the lowest id).  Any hit:
with the default (non-speculative) leaf order every ray walks the tree in the
oracle's order, so the first hit found is the oracle's, bit for bit; with the
speculative order hit/miss is identical and the reported (id, t) is a real
hit of that ray.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import random_rays, random_triangles

pytestmark = pytest.mark.gpu

# lds_nodes: how many top-of-tree node records are read from LDS (0 = none)
CONFIGS = [
    dict(block=128, lds_stack=16, chunk=64, fetch_threshold=24, waves_per_cu=16, speculative=0, leaf_gate=0,
         lds_nodes=0),
    dict(block=64, lds_stack=8, chunk=256, fetch_threshold=0, waves_per_cu=8, speculative=0, leaf_gate=40,
         lds_nodes=128),
    dict(block=256, lds_stack=24, chunk=64, fetch_threshold=48, waves_per_cu=32, speculative=0, leaf_gate=16,
         lds_nodes=512),
    dict(block=256, lds_stack=16, chunk=64, fetch_threshold=40, waves_per_cu=32, speculative=1, leaf_gate=0,
         lds_nodes=256),
    dict(block=1024, lds_stack=8, chunk=128, fetch_threshold=48, waves_per_cu=32, speculative=0, leaf_gate=16,
         lds_nodes=512),
]


def _compact2(rbm, bvh):
    from renderbox2_amd import _lib as L
    n, w, x = bvh.export(L.RB_LAYOUT_REF_COMPACT2)
    return n.reshape(-1, 4), w.reshape(-1, 4), x


@pytest.fixture(scope="module")
def small(gpu):
    v, idx = random_triangles(3000, seed=50)
    b = gpu.BVH((v, idx))
    return v, idx, b, gpu.CudaTracer(b), _compact2(gpu, b)


@pytest.fixture(scope="module")
def room(gpu):
    s = gpu.scenes.ao_room()
    b = gpu.scenes.build(s, max_leaf=4)
    return s, b, gpu.CudaTracer(b, s), _compact2(gpu, b)


def _check_any(orc, geom, rays, ids, ts, aid, at):
    assert np.array_equal(aid >= 0, ids >= 0)
    h = aid >= 0
    assert np.all((at[h] > rays[h, 3]) & (at[h] < rays[h, 7]) & (at[h] >= ts[h]))
    # the any-hit (id, t) is a real hit: the Woop test of triangle id on that
    # ray gives exactly t, inside the triangle
    v, idx = geom
    t1, ok = orc.woop_hits(v, idx, aid[h], rays[h])
    assert ok.all()
    assert np.array_equal(t1.view(np.uint32), at[h].view(np.uint32))


@pytest.mark.parametrize("wide", [1, 0])
@pytest.mark.parametrize("cfg", range(len(CONFIGS)))
def test_trace_small_bit_exact(gpu, orc, small, cfg, wide):
    """Closest and any hit of every launch shape equal the oracle; wide=1
    walks the 4-child records for closest hits, wide=0 the binary ones."""
    v, idx, b, tr, c2 = small
    saved = gpu.get_trace_config()
    gpu.set_trace_config(**CONFIGS[cfg], wide=wide)
    try:
        rays = random_rays(50000, seed=60 + cfg)
        ids, ts = tr.trace(rays)
        oid, ot = orc.trace_compact2(*c2, rays)
        assert (ids >= 0).sum() > 10000
        assert np.array_equal(ids, oid)
        assert np.array_equal(ts.view(np.uint32), ot.view(np.uint32))
        aid, at = tr.trace(rays, any_hit=True)
        _check_any(orc, (v, idx), rays, ids, ts, aid, at)
        if not CONFIGS[cfg]["speculative"]:
            oaid, oat = orc.trace_compact2(*c2, rays, any_hit=True)
            assert np.array_equal(aid, oaid)
            assert np.array_equal(at.view(np.uint32), oat.view(np.uint32))
    finally:
        gpu.set_trace_config(**saved)


def test_reciprocal_equals_the_division_for_every_input(gpu):
    # the traversal's 1/x (hardware reciprocal + one fused Newton step when
    # that is normal, else the division) must be the IEEE quotient for all
    # 2^32 inputs: the bit-exact (id, t) of every other test rests on it
    fast, bad = gpu.selftest_rcp()
    assert bad == 0
    assert fast == 4_227_858_434  # every normal |x| < 2^126, and x = +-2^126


def test_visit_counts_equal_the_oracle(gpu, orc):
    """In the default leaf order each ray walks the tree in the oracle's
    order, so the kernel's own per-ray counts of inner nodes, triangles and
    leaves (closest hit) equal the oracle's, whatever the launch shape.
    Leaves of at most 16 references keep the MI355X records one-to-one with
    the Compact2 nodes."""
    v, idx = random_triangles(4000, seed=52)
    p = gpu.Platform()
    p.set_leaf_preferences(1, 16)
    b = gpu.BVH((v, idx), p)
    tr = gpu.CudaTracer(b)
    rays = random_rays(30000, seed=71)
    oid, ot, ocnt = orc.trace_compact2(*_compact2(gpu, b), rays, counts=True)
    assert (oid >= 0).sum() > 5000
    saved = gpu.get_trace_config()
    try:
        for cfg in CONFIGS[:3]:
            gpu.set_trace_config(**cfg)
            ids, ts, cnt = tr.trace_counts(rays)
            assert np.array_equal(ids, oid) and np.array_equal(ts.view(np.uint32), ot.view(np.uint32))
            assert np.array_equal(cnt, ocnt), cfg
    finally:
        gpu.set_trace_config(**saved)


def test_wide_walk_counts(gpu, orc, small):
    """The wide walk's own visit counts (RB_WIDE_COUNTS): the same (id, t)
    as the oracle, the root record read by every ray, and fewer records per
    ray than the binary walk reads; any hit has no wide counts."""
    v, idx, b, tr, c2 = small
    rays = random_rays(30000, seed=81)
    oid, ot = orc.trace_compact2(*c2, rays)
    ids, ts, wc = tr.trace_counts(rays, wide=True)
    assert np.array_equal(ids, oid) and np.array_equal(ts.view(np.uint32), ot.view(np.uint32))
    bid, _, bc = tr.trace_counts(rays)
    assert np.array_equal(bid, oid)
    assert (wc[:, 0] >= 1).all()
    assert wc[:, 0].sum() < 0.75 * bc[:, 0].sum()
    assert (wc[ids >= 0, 1] >= 1).all() and (wc[ids >= 0, 2] >= 1).all()  # a hit tested its triangle
    with pytest.raises(gpu.RbError):
        tr.trace_counts(rays[:64], any_hit=True, wide=True)


def test_unknown_trace_flags_are_refused(gpu, small):
    # bits beyond RB_ANY / RB_MASKED (and RB_WIDE_COUNTS for the counts
    # call) are refused before any work, the library's internal ones included
    from renderbox2_amd import _lib as L
    v, idx, b, tr, c2 = small
    rays = gpu.DeviceBuffer.from_array(random_rays(64, seed=3))
    hits = gpu.DeviceBuffer(64 * 8)
    counts = gpu.DeviceBuffer(64 * 12)
    for bit in (L.RB_WIDE_COUNTS, 1 << 16, 1 << 31):
        assert L.lib().rb_trace(tr.handle, rays.ptr, 64, bit, None, hits.ptr, None) == L.RB_ERR_INVALID
    assert L.lib().rb_trace_counts(tr.handle, rays.ptr, 64, 1 << 16, hits.ptr, counts.ptr, None) == L.RB_ERR_INVALID
    assert L.lib().rb_trace(tr.handle, rays.ptr, 64, L.RB_ANY, None, hits.ptr, None) == L.RB_OK
    assert len(hits.download(np.int32, 128)) == 128  # waits for the trace


def test_trace_room_camera_and_diffuse_rays(gpu, orc, room):
    s, b, tr, c2 = room
    cam = gpu.scenes.ao_camera(256, 256)
    n = cam.get_total_pixels()
    d_prim = gpu.DeviceBuffer(n * 32)
    cam.generate_rays_device(d_prim, spp=1, seed=3, jitter=True)
    prim = d_prim.download(np.float32, n * 8).reshape(n, 8)
    # camera rays equal the oracle's restatement bit for bit
    lib = orc.lib()
    P = C.c_void_p
    lib.orc_raygen_camera.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint64, C.c_int, P]
    r2c, c2w = cam.matrices()
    r2c = np.ascontiguousarray(r2c)
    c2w = np.ascontiguousarray(c2w)
    orays = np.empty((n, 8), np.float32)
    lib.orc_raygen_camera(r2c.ctypes.data_as(P), c2w.ctypes.data_as(P), 256, 0, 0, 256, 256, 1, 3, 1,
                          orays.ctypes.data_as(P))
    assert np.array_equal(prim.view(np.uint32), orays.view(np.uint32))
    d_hits = gpu.DeviceBuffer(n * 8)
    tr.trace_device(d_prim, n, d_hits)
    hits = d_hits.download(gpu.HIT_DTYPE, n)
    oid, ot = orc.trace_compact2(*c2, prim)
    assert np.array_equal(hits["id"], oid) and np.array_equal(hits["t"].view(np.uint32), ot.view(np.uint32))
    assert (oid >= 0).all()  # closed room
    # diffuse rays from the hits
    nd = 2 * n
    d_diff = gpu.DeviceBuffer(nd * 32)
    gpu.raygen_diffuse_device(tr, d_prim, d_hits, n, nd, 11, d_diff)
    diff = d_diff.download(np.float32, nd * 8).reshape(nd, 8)
    v, idx, _, _ = s.geometry()
    lib.orc_raygen_diffuse.argtypes = [P, P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint64, P]
    od = np.empty((nd, 8), np.float32)
    lib.orc_raygen_diffuse(v.ctypes.data_as(P), idx.ctypes.data_as(P), prim.ctypes.data_as(P),
                           oid.ctypes.data_as(P), ot.ctypes.data_as(P), n, nd, 11, od.ctypes.data_as(P))
    assert np.array_equal(diff.view(np.uint32), od.view(np.uint32))
    ids, ts = tr.trace(diff)
    oid2, ot2 = orc.trace_compact2(*c2, diff)
    assert np.array_equal(ids, oid2) and np.array_equal(ts.view(np.uint32), ot2.view(np.uint32))
    aid, at = tr.trace(diff, any_hit=True)
    _check_any(orc, (v, idx), diff, ids, ts, aid, at)


def test_masked_degenerate_and_empty_batches(gpu, orc, small):
    v, idx, b, tr, c2 = small
    rays = random_rays(4096, seed=70)
    rays[0, 4:7] = 0.0                       # zero direction
    rays[1, 4:7] = np.nan                    # NaN direction
    rays[2, 3], rays[2, 7] = 5.0, 1.0        # tmin > tmax
    rays[3, 4:7] = [0.0, 0.0, 1.0]           # axis-aligned
    rays[4, 0:3] = np.nan                    # NaN origin
    mask = (np.arange(4096) % 3 != 0).astype(np.uint8)
    ids, ts = tr.trace(rays, mask=mask)
    oid, ot = orc.trace_compact2(*c2, rays)
    m = mask.astype(bool)
    assert np.array_equal(ids[m], oid[m]) and np.array_equal(ts[m].view(np.uint32), ot[m].view(np.uint32))
    assert np.all(ids[~m] == -1) and np.array_equal(ts[~m], rays[~m, 7])
    assert ids[1] == -1 and ids[2] == -1 and ids[4] == -1
    # empty and single-ray batches
    e_ids, e_ts = tr.trace(rays[:0])
    assert e_ids.size == 0
    one_id, one_t = tr.trace(rays[5:6])
    assert one_id[0] == oid[5] and one_t[0] == ot[5]


def _flat_quads(n, seed):
    # axis-aligned quads: every box is flat along one axis (zero extent)
    rng = np.random.default_rng(seed)
    v, idx = [], []
    for i in range(n):
        a = i % 3
        c = rng.uniform(-5, 5, 3)
        u, w = [k for k in range(3) if k != a]
        corners = []
        for du, dw in ((0, 0), (1, 0), (1, 1), (0, 1)):
            q = c.copy()
            q[u] += du * 0.5
            q[w] += dw * 0.5
            corners.append(q)
        base = len(v)
        v += corners
        idx += [[base, base + 1, base + 2], [base, base + 2, base + 3]]
    return np.array(v, np.float32), np.array(idx, np.uint32)


@pytest.mark.parametrize("case", ["one_triangle", "flat_quads", "tiny_coords", "large_coords", "huge_coords"])
def test_wide_walk_edge_scenes_equal_the_oracle(gpu, orc, case):
    """Scenes at the edges of the 8-bit record encoding, closest hit with
    the wide records against the oracle: a single triangle (a root leaf
    beside an empty one), zero-thickness boxes, coordinates scaled by 1e-9
    and 1e9 (step exponents near -40 and +20), and by 1e30, where the
    builder's SAH overflows and leaves inverted boxes (held as the box
    between their planes; the Woop rows overflow too, so every ray misses
    in both walks)."""
    if case == "one_triangle":
        v = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
        idx = np.array([[0, 1, 2]], np.uint32)
        rays = random_rays(4096, seed=11, box=2.0)
    elif case == "flat_quads":
        v, idx = _flat_quads(1500, 12)
        rays = random_rays(20000, seed=12)
    else:
        sc = np.float32({"tiny_coords": 1e-9, "large_coords": 1e9, "huge_coords": 1e30}[case])
        v, idx = random_triangles(2000, seed=13)
        v = (v * sc).astype(np.float32)
        rays = random_rays(20000, seed=14)
        rays[:, 0:3] *= sc
        rays[:, 7] = 3.0e38
    b = gpu.BVH((v, idx))
    tr = gpu.CudaTracer(b)
    assert tr.device_view().wnodes  # the wide walk is what runs
    ids, ts = tr.trace(rays)
    oid, ot = orc.trace_compact2(*_compact2(gpu, b), rays)
    assert (oid >= 0).sum() > (-1 if case == "huge_coords" else 20)
    assert np.array_equal(ids, oid) and np.array_equal(ts.view(np.uint32), ot.view(np.uint32))


@pytest.mark.parametrize("n", [1, 7, 64, 65, 511, 513, 4097])
def test_ray_band_edges_equal_the_oracle(gpu, orc, small, n):
    """Launch sizes around the ray bands' edges (a launch's rays are split
    into eight bands of whole 64-ray chunks, the last one ending at the ray
    count): every ray is traced once, closest and any hit."""
    v, idx, b, tr, c2 = small
    rays = random_rays(n, seed=200 + n)
    ids, ts = tr.trace(rays)
    oid, ot = orc.trace_compact2(*c2, rays)
    assert np.array_equal(ids, oid) and np.array_equal(ts.view(np.uint32), ot.view(np.uint32))
    aid, at = tr.trace(rays, any_hit=True)
    oaid, oat = orc.trace_compact2(*c2, rays, any_hit=True)
    assert np.array_equal(aid, oaid) and np.array_equal(at.view(np.uint32), oat.view(np.uint32))


def test_trace_is_deterministic_and_reentrant(gpu, small):
    v, idx, b, tr, c2 = small
    rays = random_rays(100000, seed=80)
    dr = gpu.DeviceBuffer.from_array(rays)
    outs = []
    streams = [gpu.Stream(), gpu.Stream()]
    bufs = [gpu.DeviceBuffer(len(rays) * 8) for _ in range(4)]
    for k in range(4):   # concurrent launches on two streams share nothing
        tr.trace_device(dr, len(rays), bufs[k], stream=streams[k % 2])
    for s in streams:
        s.synchronize()
    for k in range(4):
        outs.append(bufs[k].download(np.uint32, 2 * len(rays)))
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


def test_trace_fill_matches_hits(gpu, room):
    s, b, tr, c2 = room
    from renderbox2_amd import _lib as L
    rays = random_rays(20000, seed=90, box=4.0)
    n = len(rays)
    dr = gpu.DeviceBuffer.from_array(rays)
    dh = gpu.DeviceBuffer(n * 8)
    bufs = {k: gpu.DeviceBuffer(n * 4) for k in ("px", "py", "pz", "nx", "ny", "nz", "pid", "mid", "eps", "t")}
    inter = gpu.DeviceBuffer(n)
    soa = L.rb_isect_soa()
    soa.position[0], soa.position[1], soa.position[2] = bufs["px"].ptr, bufs["py"].ptr, bufs["pz"].ptr
    soa.geometric_normal[0], soa.geometric_normal[1], soa.geometric_normal[2] = bufs["nx"].ptr, bufs["ny"].ptr, bufs["nz"].ptr
    soa.primitive_id, soa.material_id, soa.epsilon, soa.t = bufs["pid"].ptr, bufs["mid"].ptr, bufs["eps"].ptr, bufs["t"].ptr
    soa.intersected = inter.ptr
    L.check(L.lib().rb_trace_fill(tr.handle, dr.ptr, n, 0, None, dh.ptr, C.byref(soa), None))
    hits = dh.download(gpu.HIT_DTYPE, n)
    got = {k: bufs[k].download(np.float32 if k not in ("pid", "mid") else np.uint32, n) for k in bufs}
    it = inter.download(np.uint8, n)
    h = hits["id"] >= 0
    assert np.array_equal(it.astype(bool), h)
    assert np.array_equal(got["pid"][h], hits["id"][h].astype(np.uint32))
    assert np.array_equal(got["t"], hits["t"])
    p = rays[h, :3] + rays[h, 4:7] * hits["t"][h, None]
    assert np.allclose(np.stack([got["px"][h], got["py"][h], got["pz"][h]], 1), p, atol=1e-4)
    v, idx, _, _ = s.geometry()
    tri = v[idx[hits["id"][h]]]
    ng = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    assert np.allclose(np.stack([got["nx"][h], got["ny"][h], got["nz"][h]], 1), ng, atol=1e-5)
    assert np.allclose(got["eps"][h], 1e-3 * hits["t"][h])


def test_million_triangles_properties(gpu, orc):
    """Full-size scene (1M triangles): a 64k-ray sample bit-exact with the oracle,
    any-hit <=> closest-hit on all 1M rays, and run-to-run determinism."""
    s = gpu.scenes.trace_1m()
    b = gpu.scenes.build(s)
    tr = gpu.CudaTracer(b, s)
    cam = gpu.scenes.trace_camera(1024, 512)
    n = cam.get_total_pixels()
    d_prim = gpu.DeviceBuffer(n * 32)
    cam.generate_rays_device(d_prim, seed=1)
    d_hits = gpu.DeviceBuffer(n * 8)
    tr.trace_device(d_prim, n, d_hits)
    d_diff = gpu.DeviceBuffer(2 * n * 32)
    gpu.raygen_diffuse_device(tr, d_prim, d_hits, n, 2 * n, 5, d_diff)
    diff = d_diff.download(np.float32, 2 * n * 8).reshape(-1, 8)
    ids, ts = tr.trace(diff)
    ids2, ts2 = tr.trace(diff)
    assert np.array_equal(ids, ids2) and np.array_equal(ts, ts2)
    aid, _ = tr.trace(diff, any_hit=True)
    assert np.array_equal(aid >= 0, ids >= 0)
    c2 = _compact2(gpu, b)
    sel = np.random.default_rng(0).choice(len(diff), 65536, replace=False)
    oid, ot = orc.trace_compact2(*c2, diff[sel])
    assert np.array_equal(ids[sel], oid) and np.array_equal(ts[sel].view(np.uint32), ot.view(np.uint32))
    oaid, _ = orc.trace_compact2(*c2, diff[sel], any_hit=True)
    assert np.array_equal(aid[sel], oaid)
    # every launch shape and both record widths give the same answers: each
    # ray's visit order is its own, and ties at equal t go to the lowest id
    saved = gpu.get_trace_config()
    try:
        for cfg in [dict(saved, wide=1 - saved["wide"])] + CONFIGS[1:3]:
            gpu.set_trace_config(**cfg)
            i3, t3 = tr.trace(diff)
            a3, _ = tr.trace(diff, any_hit=True)
            assert np.array_equal(i3, ids) and np.array_equal(t3.view(np.uint32), ts.view(np.uint32))
            assert np.array_equal(a3, aid)
    finally:
        gpu.set_trace_config(**saved)


def test_batched_trace_equals_the_oracle_per_batch(gpu, orc):
    """Uneven batches against several resident scenes, closest and any hit,
    more batches than one launch takes (64), empty batches: every batch of
    one batched call (the kernel's batch-table path: a lane finds its ray's
    batch and scene when it claims the ray) equals the oracle on that batch's
    scene, (id, t) bit for bit -- any hit too, since each ray walks its tree
    in the oracle's order and stops at the same first hit."""
    rng = np.random.default_rng(3)
    scenes = []
    for k, n in enumerate([800, 3000, 20000]):
        v, idx = random_triangles(n, seed=100 + k, spread=4.0 + k)
        b = gpu.BVH((v, idx))
        scenes.append((gpu.CudaTracer(b), _compact2(gpu, b)))
    batches, inputs = [], []
    for i in range(70):
        n = 0 if i % 17 == 5 else int(np.exp(rng.uniform(np.log(64), np.log(40000))))
        tr, c2 = scenes[i % 3]
        anyh = i % 4 == 3
        rays = random_rays(max(n, 1), seed=1000 + i, box=7.0)[:n]
        dr = gpu.DeviceBuffer.from_array(rays) if n else gpu.DeviceBuffer(32)
        dh = gpu.DeviceBuffer(max(n, 1) * 8)
        batches.append((tr, dr, n, dh, anyh))
        inputs.append((rays, c2))
    gpu.trace_batch(batches)
    checked_any = 0
    for (tr, dr, n, dh, anyh), (rays, c2) in zip(batches, inputs):
        if not n:
            continue
        h = dh.download(gpu.HIT_DTYPE, n)
        oid, ot = orc.trace_compact2(*c2, rays, any_hit=anyh)
        assert np.array_equal(h["id"], oid) and np.array_equal(h["t"].view(np.uint32), ot.view(np.uint32)), anyh
        checked_any += anyh
    assert checked_any >= 15


def test_device_function_trace_equals_the_oracle(gpu, orc, small, room):
    """rb_trace_one (the header-only device tracer, the reference's trace_ray
    form: one ray per thread, cuda_tracer_functions.cu:42-279) against the
    oracle: closest and any hit, (id, t) bit for bit."""
    v, idx, b, tr, c2 = small
    rays = random_rays(20000, seed=77)
    for any_hit in (False, True):
        s = tr.trace_simple(rays, any_hit=any_hit)
        oid, ot = orc.trace_compact2(*c2, rays, any_hit=any_hit)
        assert np.array_equal(s[0], oid) and np.array_equal(s[1].view(np.uint32), ot.view(np.uint32)), any_hit
    sc, bb, rt, rc2 = room
    cam = gpu.scenes.ao_camera(128, 128)
    d = gpu.DeviceBuffer(cam.get_total_pixels() * 32)
    cam.generate_rays_device(d, seed=3)
    prim = d.download(np.float32, cam.get_total_pixels() * 8).reshape(-1, 8)
    for any_hit in (False, True):
        s = rt.trace_simple(prim, any_hit=any_hit)
        oid, ot = orc.trace_compact2(*rc2, prim, any_hit=any_hit)
        assert np.array_equal(s[0], oid) and np.array_equal(s[1].view(np.uint32), ot.view(np.uint32)), any_hit
    view = rt.device_view()
    assert view.nodes and view.tris and view.index and view.n_refs >= view.n_nodes // 2 and view.errors


def test_trace_matches_the_golden_fixture(gpu):
    """The HIP path on the seeded 2k-triangle scene against the committed
    fixture: closest and any-hit (id, t) and per-ray visit counts."""
    import importlib.util
    import os
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    spec = importlib.util.spec_from_file_location("make_fixtures", os.path.join(golden, "make_fixtures.py"))
    mf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mf)
    g = np.load(os.path.join(golden, "trace_2k.npz"))
    v, idx, rays = mf.trace_2k_inputs()
    p = gpu.Platform()
    p.set_leaf_preferences(1, mf.SCENE["max_leaf"])
    tr = gpu.CudaTracer(gpu.BVH((v, idx), p))
    ids, ts, cnt = tr.trace_counts(rays)
    assert np.array_equal(ids, g["ids"]) and np.array_equal(ts.view(np.uint32), g["t"].view(np.uint32))
    assert np.array_equal(cnt, g["counts"])
    aid, at = tr.trace(rays, any_hit=True)
    assert np.array_equal(aid, g["any_ids"]) and np.array_equal(at.view(np.uint32), g["any_t"].view(np.uint32))


def test_largest_batch_one_call_takes(gpu, small):
    """2^31 - 1 rays in one rb_trace call (68.7 GB of rays, 17.2 GB of hits,
    every ray and hit offset past 2^31 bytes exercised): a 2^20-ray set tiled
    over the whole batch, so every hit must equal the hit of its ray in the
    small set, traced alone."""
    from renderbox2_amd import _lib as L
    v, idx, b, tr, c2 = small
    nb = 1 << 20
    base = random_rays(nb, seed=91)
    bid, bt = tr.trace(base)
    n = 0x7FFFFFFF
    d_rays = gpu.DeviceBuffer(n * 32)
    d_rays.upload(base)
    filled = nb
    while filled < n:  # doubling copies of the filled prefix
        k = min(filled, n - filled)
        L.check(L.lib().rb_memcpy_dtod(d_rays.ptr + filled * 32, d_rays.ptr, k * 32, None))
        filled += k
    d_hits = gpu.DeviceBuffer(n * 8)
    tr.trace_device(d_rays, n, d_hits)
    L.check(L.lib().rb_stream_synchronize(None))
    d_rays.free()
    chunk = 1 << 24
    exp_id, exp_t = np.tile(bid, chunk // nb), np.tile(bt.view(np.uint32), chunk // nb)
    for s in range(0, n, chunk):
        cnt = min(chunk, n - s)
        out = np.empty(cnt, gpu.HIT_DTYPE)
        L.check(L.lib().rb_memcpy_dtoh(out.ctypes.data, d_hits.ptr + s * 8, cnt * 8, None))
        assert np.array_equal(out["id"], exp_id[:cnt]), s
        assert np.array_equal(out["t"].view(np.uint32), exp_t[:cnt]), s
    d_hits.free()


def test_counter_slots_reused_across_streams(gpu, orc):
    """A launch's work-counter pair comes from a ring of 4,096; a slot is
    reused 4,096 launches later, and when that launch is on another stream it
    first waits for the slot's previous launch.  One long launch on stream A
    (few waves per CU, so other streams' work runs beside it) and 4,200
    launches on stream B: every hit of both equals the oracle's."""
    v, idx = random_triangles(60000, seed=120, spread=8.0)
    b = gpu.BVH((v, idx))
    tr = gpu.CudaTracer(b)
    c2 = _compact2(gpu, b)
    a_rays = random_rays(1 << 22, seed=121, box=9.0)
    b_rays = random_rays(4200 * 256, seed=122, box=9.0)
    d_a, d_ah = gpu.DeviceBuffer.from_array(a_rays), gpu.DeviceBuffer(len(a_rays) * 8)
    d_b, d_bh = gpu.DeviceBuffer.from_array(b_rays), gpu.DeviceBuffer(len(b_rays) * 8)
    sa, sb = gpu.Stream(), gpu.Stream()
    saved = gpu.get_trace_config()
    try:
        gpu.set_trace_config(**dict(saved, waves_per_cu=1))
        tr.trace_device(d_a, len(a_rays), d_ah, stream=sa)
        gpu.set_trace_config(**saved)
        from renderbox2_amd import _lib as L
        lib = L.lib()
        for k in range(4200):
            L.check(lib.rb_trace(tr.handle, d_b.ptr + k * 256 * 32, 256, L.RB_CLOSEST, None,
                                 d_bh.ptr + k * 256 * 8, sb.handle))
    finally:
        gpu.set_trace_config(**saved)
    sa.synchronize()
    sb.synchronize()
    ha = d_ah.download(gpu.HIT_DTYPE, len(a_rays))
    hb = d_bh.download(gpu.HIT_DTYPE, len(b_rays))
    for h, rays in ((ha, a_rays), (hb, b_rays)):
        oid, ot = orc.trace_compact2(*c2, rays)
        assert np.array_equal(h["id"], oid) and np.array_equal(h["t"].view(np.uint32), ot.view(np.uint32))


def test_no_traversal_errors_were_flagged(gpu):
    """Every BVH these tests uploaded passed the upload's depth check, so no
    traversal stack overflowed in any launch."""
    flags = C.c_uint32(0xFFFFFFFF)
    from renderbox2_amd import _lib as L
    assert L.lib().rb_trace_errors(C.byref(flags)) == L.RB_OK, L.last_error()
    assert flags.value == 0
