"""Pins of the oracle (the CPU restatement in oracle/rb_oracle.c).

No reference test, fixture or golden vector exists for this path (the
reference ships no tests), so every oracle function is pinned here by
closed forms, FP64 brute force, invariants and published known answers.
"""
import os

import numpy as np
import pytest

from conftest import random_rays, random_triangles

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# ---------------------------------------------------------------- math --


def test_invert4x4_matches_float64_inverse(orc):
    rng = np.random.default_rng(7)
    for _ in range(50):
        m = rng.normal(0, 1, (4, 4)).astype(np.float32) + 3 * np.eye(4, dtype=np.float32)
        inv = orc.invert4x4(m)
        ref = np.linalg.inv(m.astype(np.float64))
        assert np.allclose(inv, ref, rtol=1e-4, atol=1e-5)
        assert np.allclose(inv.astype(np.float64) @ m, np.eye(4), atol=1e-4)


def test_invert4x4_cofactor_order_plants_a_transpose():
    """A transposed cofactor (a plausible slip) must fail the inverse check."""
    rng = np.random.default_rng(3)
    m = rng.normal(0, 1, (4, 4)) + 3 * np.eye(4)
    assert not np.allclose(np.linalg.inv(m).T @ m, np.eye(4), atol=1e-3)


def test_woopify_unit_triangle_closed_form(orc):
    # v0=(1,0,0) v1=(0,1,0) v2=(0,0,0): M = [e0 e1 n v2] = I, so the rows are
    # (n row 2 = (0,0,1,-0)), (1,0,0,0), (0,1,0,0).
    w = orc.woopify([1, 0, 0], [0, 1, 0], [0, 0, 0])
    assert np.array_equal(w[0], np.array([0, 0, 1, 0], np.float32))
    assert np.array_equal(w[1], np.array([1, 0, 0, 0], np.float32))
    assert np.array_equal(w[2], np.array([0, 1, 0, 0], np.float32))


def test_woopify_gives_plane_distance_and_barycentrics(orc):
    rng = np.random.default_rng(11)
    for _ in range(200):
        v = rng.normal(0, 2, (3, 3)).astype(np.float32)
        w = orc.woopify(v[0], v[1], v[2]).astype(np.float64)
        # a point p = a v0 + b v1 + (1-a-b) v2 maps to (u, v) = (a, b) and
        # lies on the plane (Oz row gives 0).
        a, b = rng.uniform(0, 0.5, 2)
        p = a * v[0] + b * v[1] + (1 - a - b) * v[2]
        u = w[1, :3] @ p + w[1, 3]
        vv = w[2, :3] @ p + w[2, 3]
        z = w[0, :3] @ p - w[0, 3]
        scale = np.abs(v).max()
        assert abs(u - a) < 2e-4 * scale and abs(vv - b) < 2e-4 * scale
        assert abs(z) < 2e-4 * scale


def test_philox_known_answers(orc):
    # Random123 kat_vectors for philox4x32_10 (Salmon et al., SC'11), tests/golden/philox4x32_10_kat.json
    import json
    with open(os.path.join(GOLDEN, "philox4x32_10_kat.json")) as f:
        kat = json.load(f)
    assert len(kat) == 3
    for v in kat:
        assert list(orc.philox(v["counter"], v["key"])) == v["out"]


def _golden_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_fixtures", os.path.join(GOLDEN, "make_fixtures.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_oracle_reproduces_the_golden_trace_fixture(orc):
    """The oracle's SBVH, Compact2 bytes and traversal (ids, t, visit counts)
    on the seeded 2k-triangle scene equal the committed fixture."""
    import hashlib
    mf = _golden_module()
    g = np.load(os.path.join(GOLDEN, "trace_2k.npz"))
    v, idx, rays = mf.trace_2k_inputs()
    b = orc.Bvh(v, idx, max_leaf=mf.SCENE["max_leaf"])
    nodes, woop, tidx = b.compact2()
    assert [nodes.size, woop.size, tidx.size] == g["sizes"].tolist()
    for name, a in (("sha_nodes", nodes), ("sha_woop", woop), ("sha_index", tidx)):
        assert hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest() == g[name].tobytes(), name
    st = b.stats()
    assert [st[k] for k in ("n_inner", "n_leaf", "n_tris", "n_child_nodes", "duplicates", "max_depth")] == \
        g["stats"].tolist()
    ids, t, cnt = orc.trace_compact2(nodes, woop, tidx, rays, counts=True)
    assert np.array_equal(ids, g["ids"]) and np.array_equal(t.view(np.uint32), g["t"].view(np.uint32))
    assert np.array_equal(cnt, g["counts"])
    aid, at = orc.trace_compact2(nodes, woop, tidx, rays, any_hit=True)
    assert np.array_equal(aid, g["any_ids"]) and np.array_equal(at.view(np.uint32), g["any_t"].view(np.uint32))


def test_sincos_2pi_against_float64(orc):
    import ctypes as C
    lib = orc.lib()
    lib.orc_sincos_2pi.argtypes = [C.c_float, C.c_void_p, C.c_void_p]
    s, c = C.c_float(), C.c_float()
    worst = 0.0
    for u in np.linspace(0, 1, 4097, endpoint=False, dtype=np.float32):
        lib.orc_sincos_2pi(float(u), C.byref(s), C.byref(c))
        worst = max(worst, abs(s.value - np.sin(2 * np.pi * float(u))), abs(c.value - np.cos(2 * np.pi * float(u))))
    assert worst < 3e-7


def test_cosine_direction_is_unit_and_above_surface(orc):
    import ctypes as C
    lib = orc.lib()
    lib.orc_cosine_dir.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p]
    rng = np.random.default_rng(5)
    cos_sum = 0.0
    N = 4000
    for _ in range(N):
        n = rng.normal(0, 1, 3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        out = np.zeros(3, np.float32)
        lib.orc_cosine_dir(n.ctypes.data_as(C.c_void_p), float(rng.uniform()), float(rng.uniform()),
                           out.ctypes.data_as(C.c_void_p))
        assert abs(np.linalg.norm(out) - 1) < 1e-5
        cth = float(out @ n)
        assert cth >= -1e-6
        cos_sum += cth
    # E[cos theta] = 2/3 for a cosine-weighted hemisphere
    assert abs(cos_sum / N - 2.0 / 3.0) < 0.02


# ---------------------------------------------------------------- SBVH --


def _decode_compact2(nodes, woop, tidx):
    """Walk Compact2 rows: returns the list of leaves (tri ids) and checks boxes nest."""
    leaves = []
    stack = [0]
    while stack:
        r = stack.pop()
        n = nodes[r:r + 4]
        b = n[:3].view(np.float32)
        boxes = [(b[0, 0], b[0, 1], b[0, 2], b[0, 3], b[2, 0], b[2, 1]),
                 (b[1, 0], b[1, 1], b[1, 2], b[1, 3], b[2, 2], b[2, 3])]
        for k in range(2):
            c = np.int32(n[3, k])
            if c >= 0:
                stack.append(int(c))
            else:
                row = ~int(c)
                ids = []
                while woop[row, 0] != 0x80000000:
                    ids.append(int(tidx[row]))
                    row += 3
                leaves.append((ids, boxes[k]))
    return leaves


def test_sbvh_invariants(orc):
    v, idx = random_triangles(2500, seed=2)
    b = orc.Bvh(v, idx)
    st = b.stats()
    info, bnd = b.nodes()
    tris = b.tri_indices()
    # every triangle is referenced (none of these is degenerate)
    assert set(tris.tolist()) == set(range(len(idx)))
    assert st["n_tris"] == len(tris) == len(idx) + st["duplicates"]
    assert st["n_inner"] + 1 == st["n_leaf"]
    # child boxes inside parent boxes; leaf boxes inside their triangles' bounds
    for k in range(len(info)):
        leaf, c0, c1, lo, hi = info[k]
        if not leaf:
            for c in (c0, c1):
                assert np.all(bnd[c, :3] >= bnd[k, :3]) and np.all(bnd[c, 3:] <= bnd[k, 3:])
        else:
            for t in tris[lo:hi]:
                tv = v[idx[t]]
                # the leaf box overlaps the triangle's box (it is a clipped part of it)
                assert np.all(bnd[k, :3] <= tv.max(0) + 1e-6) and np.all(bnd[k, 3:] >= tv.min(0) - 1e-6)
    # SAH recomputed in float64 from the dump agrees with the builder's value
    def area(bb):
        d = np.maximum(bb[3:] - bb[:3], 0)
        return 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
    sah = 0.0
    prob = {0: 1.0}
    for k in range(len(info)):
        leaf, c0, c1, lo, hi = info[k]
        sah += prob[k] * (2.0 if not leaf else float(hi - lo))
        if not leaf:
            for c in (c0, c1):
                prob[c] = prob[k] * area(bnd[c]) / area(bnd[k])
    assert abs(sah - st["sah"]) < 1e-3 * sah


def test_sbvh_tiny_scene_by_hand(orc):
    # Two far-apart unit triangles in z = 0.  Root box 11 x 1 x 0: area 22.
    # Costs 1/1: leafSAH = 2*22 = 44 = nodeSAH alone, so the root stays a leaf.
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [10, 0, 0], [11, 0, 0], [10, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2], [3, 4, 5]], np.uint32)
    st = orc.Bvh(v, idx).stats()
    assert (st["n_inner"], st["n_leaf"]) == (0, 1)
    assert abs(st["sah"] - 2.0) < 1e-6  # probability 1 x 2 triangles
    # Node cost 0.1: split SAH = 0.2*22 + 2*1 + 2*1 = 8.4 < 44, two leaves of one.
    b = orc.Bvh(v, idx, node_cost=0.1)
    st = b.stats()
    assert (st["n_inner"], st["n_leaf"], st["duplicates"]) == (1, 2, 0)
    nodes, woop, tidx = b.compact2()
    assert nodes.shape == (4, 4)
    # child 0 = left (lower x) leaf = triangle 0; both children are leaves
    c0, c1 = np.int32(nodes[3, 0]), np.int32(nodes[3, 1])
    assert c0 < 0 and c1 < 0
    assert tidx[~int(c0)] == 0 and tidx[~int(c1)] == 1
    bx = nodes[:3].view(np.float32)
    assert (bx[0, 0], bx[0, 1]) == (0.0, 1.0) and (bx[1, 0], bx[1, 1]) == (10.0, 11.0)


def test_sbvh_degenerate_triangles_dropped(orc):
    # A line (zero area, two zero extents) is removed by the degenerate test.
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [2, 0, 0], [3, 0, 0], [4, 0, 0], [5, 5, 5], [6, 5, 5], [5, 6, 5]],
                 np.float32)
    idx = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 8]], np.uint32)
    b = orc.Bvh(v, idx)
    assert sorted(b.tri_indices().tolist()) == [0, 2]


def test_compact2_decodes_to_the_tree(orc):
    v, idx = random_triangles(800, seed=4)
    b = orc.Bvh(v, idx, max_leaf=4)
    nodes, woop, tidx = b.compact2()
    leaves = _decode_compact2(nodes, woop, tidx)
    info, bnd = b.nodes()
    tris = b.tri_indices()
    want = sorted(sorted(tris[lo:hi].tolist()) for leaf, _, _, lo, hi in info if leaf)
    got = sorted(sorted(ids) for ids, _ in leaves)
    assert got == want
    assert all(len(ids) <= 4 for ids, _ in leaves)
    # terminators: one per leaf; no Woop row starts with -0
    term = np.where(woop[:, 0] == 0x80000000)[0]
    assert len(term) == len(leaves)


# ------------------------------------------------------------ traversal --


def _ambiguous(margin, gap):
    # grazing hits (within 1e-5 of an edge) and near-ties (another triangle
    # within 1e-5 relative in t) may legitimately differ between FP32 Woop and FP64.
    return (margin < 1e-5) | (gap < 1e-5)


@pytest.mark.parametrize("seed", [0, 1])
def test_trace_matches_fp64_bruteforce(orc, seed):
    v, idx = random_triangles(3000, seed=10 + seed)
    b = orc.Bvh(v, idx)
    nodes, woop, tidx = b.compact2()
    rays = random_rays(20000, seed=20 + seed)
    ids, ts = orc.trace_compact2(nodes, woop, tidx, rays)
    bid, bt, margin, gap = orc.bruteforce_f64(v, idx, rays)
    amb = _ambiguous(margin, gap)
    assert amb.mean() < 0.01
    ok = ~amb
    assert np.array_equal(ids[ok], bid[ok])
    hit = ok & (bid >= 0)
    assert hit.sum() > 5000
    err = np.abs(ts[hit] - bt[hit])
    assert np.all(err <= 1e-4 * np.maximum(bt[hit], 1.0))
    assert np.all(ts[bid < 0] == rays[bid < 0, 7])


def test_trace_any_hit_agrees_with_closest(orc):
    v, idx = random_triangles(2000, seed=30)
    b = orc.Bvh(v, idx)
    nodes, woop, tidx = b.compact2()
    rays = random_rays(10000, seed=31)
    ids, ts = orc.trace_compact2(nodes, woop, tidx, rays)
    aid, at = orc.trace_compact2(nodes, woop, tidx, rays, any_hit=True)
    assert np.array_equal(aid >= 0, ids >= 0)
    h = aid >= 0
    assert np.all((at[h] > rays[h, 3]) & (at[h] < rays[h, 7]))
    assert np.all(at[h] >= ts[h])


def test_trace_respects_tmin_tmax(orc):
    v, idx = random_triangles(2000, seed=32)
    b = orc.Bvh(v, idx)
    nodes, woop, tidx = b.compact2()
    rays = random_rays(5000, seed=33)
    ids, ts = orc.trace_compact2(nodes, woop, tidx, rays)
    h = ids >= 0
    assert h.sum() > 1000
    # tmax just below the closest hit: nothing can be hit any more
    r2 = rays.copy()
    r2[h, 7] = np.nextafter(ts[h], np.float32(-np.inf))
    ids2, ts2 = orc.trace_compact2(nodes, woop, tidx, r2)
    assert np.all(ids2[h] == -1)
    assert np.array_equal(ts2[h], r2[h, 7])
    # tmin at the closest hit: the next hit is strictly further (or none)
    r3 = rays.copy()
    r3[h, 3] = ts[h]
    ids3, ts3 = orc.trace_compact2(nodes, woop, tidx, r3)
    assert np.all((ids3[h] == -1) | (ts3[h] > ts[h]))
    bid, bt, margin, gap = orc.bruteforce_f64(v, idx, r3)
    ok = h & (margin >= 1e-5) & (gap >= 1e-5) & (np.abs(bt - r3[:, 3]) > 1e-4 * np.maximum(bt, 1))
    assert np.array_equal(ids3[ok], bid[ok])


def test_tie_goes_to_lowest_id(orc):
    # Three copies of one triangle: every ray hitting it must report id 0 no
    # matter where the builder placed the copies; a shared edge hit exactly
    # picks the lower id too.
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    v = np.concatenate([base, base, base, [[5, 5, 5], [6, 5, 5], [5, 6, 5]]]).astype(np.float32)
    idx = np.arange(12, dtype=np.uint32).reshape(4, 3)
    idx = idx[[3, 2, 1, 0]].copy()  # tri 3 is the original 0 copy etc.
    b = orc.Bvh(v, idx)
    nodes, woop, tidx = b.compact2()
    n = 500
    rng = np.random.default_rng(0)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:2] = rng.uniform(0.05, 0.4, (n, 2))
    rays[:, 2] = 2.0
    rays[:, 6] = -1.0
    rays[:, 7] = 1e30
    ids, ts = orc.trace_compact2(nodes, woop, tidx, rays)
    assert np.all(ids == 1)  # copies are triangles 1, 2, 3; lowest is 1
    assert np.all(ts == 2.0)
    # two triangles sharing the edge x == y; a ray through the edge hits both at t=1
    v2 = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [9, 9, 9], [9.5, 9, 9], [9, 9.5, 9]], np.float32)
    idx2 = np.array([[4, 5, 6], [0, 2, 3], [0, 1, 2]], np.uint32)
    b2 = orc.Bvh(v2, idx2)
    n2, w2, t2 = b2.compact2()
    r = np.array([[0.5, 0.5, 1.0, 0.0, 0.0, 0.0, -1.0, 1e30]], np.float32)
    i2, tt2 = orc.trace_compact2(n2, w2, t2, r)
    assert i2[0] == 1 and tt2[0] == 1.0


def test_traversal_counts_are_consistent(orc):
    v, idx = random_triangles(3000, seed=40)
    b = orc.Bvh(v, idx)
    nodes, woop, tidx = b.compact2()
    rays = random_rays(4000, seed=41)
    ids, ts, cnt = orc.trace_compact2(nodes, woop, tidx, rays, counts=True)
    assert np.all(cnt[:, 0] >= 1)
    # every leaf reached is ended by its terminator, every tested triangle is in one
    assert np.all(cnt[:, 1] >= 0) and np.all(cnt[:, 2] <= cnt[:, 0] * 2)
    bpr = orc.bytes_per_ray(cnt, ids >= 0)
    assert np.all(bpr == 40 + 64 * cnt[:, 0] + 48 * cnt[:, 1] + 16 * cnt[:, 2] + 4 * (ids >= 0))


def test_single_triangle_woop_test_agrees_with_traversal(orc):
    v, idx = random_triangles(1500, seed=42)
    b = orc.Bvh(v, idx)
    rays = random_rays(8000, seed=43)
    ids, ts = orc.trace_compact2(*b.compact2(), rays)
    h = ids >= 0
    t, ok = orc.woop_hits(v, idx, ids[h], rays[h])
    assert ok.all() and np.array_equal(t.view(np.uint32), ts[h].view(np.uint32))
    # a wrong triangle id is (almost always) not a hit
    t2, ok2 = orc.woop_hits(v, idx, (ids[h] + 1) % len(idx), rays[h])
    assert ok2.mean() < 0.05


def _cam(cam):
    c = cam.c
    return (tuple(c.position), tuple(c.lookat), tuple(c.up), c.hfov_deg, c.width, c.height)


def _camera_rays(orc, r2c, c2w, w, h):
    return orc.raygen_camera(r2c, c2w, w, (0, 0, w, h), 1, 3)


def test_diffuse_rays_by_their_definition(orc, rbm):
    """The incoherent ray set of the trace benchmark (orc_raygen_diffuse) from
    its definition: ray i starts at the hit point o + t d of primary i, or of
    another hit primary when i missed, with tmin = 1e-3 t exactly; its
    direction is a unit vector on the side of the hit triangle's geometric
    normal that faces the primary ray; directions are cosine distributed
    (E[cos] = 2/3, E[cos^2] = 1/2, and the azimuth uniform)."""
    s = rbm.scenes.trace_scaled(20_000, seed=2)
    v, idx, _, _ = s.geometry()
    from renderbox2_amd import _lib as L
    b = rbm.BVH(s)
    n, w, x = b.export(L.RB_LAYOUT_REF_COMPACT2)
    cam = rbm.scenes.trace_camera(256, 128)
    r2c, c2w = orc.camera_matrices(*_cam(cam))
    prim = _camera_rays(orc, r2c, c2w, 256, 128)
    ids, ts = orc.trace_compact2(n.reshape(-1, 4), w.reshape(-1, 4), x, prim)
    hit = ids >= 0
    assert 0.02 < hit.mean() < 0.95  # most rays redraw their origin from another primary
    nout = 2 * len(prim)
    out = orc.raygen_diffuse(v, idx, prim, ids, ts, nout, 7)
    origins = prim[:, :3].astype(np.float64) + prim[:, 4:7] * ts[:, None].astype(np.float64)
    # every origin is the hit point of some hit primary, with that primary's tmin
    o32 = (prim[:, :3] + prim[:, 4:7] * ts[:, None]).astype(np.float32)
    own = np.arange(nout) % len(prim)
    same = hit[own]
    assert np.array_equal(out[same, :3], o32[own[same]])
    assert np.array_equal(out[same, 3], (np.float32(1e-3) * ts[own[same]]).astype(np.float32))
    # a missed primary redraws another primary up to 64 times; a ray that
    # finds none is the empty ray (zero length) -- (1 - hit fraction)^64 of them
    empty = out[:, 7] == 0
    assert np.array_equal(out[empty], np.tile(np.float32([0, 0, 0, 0, 0, 0, 1, 0]), (empty.sum(), 1)))
    assert empty.mean() < 3 * (1 - hit.mean()) ** 64 + 1e-3
    redrawn = ~same & ~empty
    keyed = {tuple(r) for r in o32[hit].view(np.uint32)}
    assert all(tuple(r) in keyed for r in out[redrawn, :3].view(np.uint32))
    tmin_of = {tuple(r): t for r, t in zip(o32[hit].view(np.uint32), (np.float32(1e-3) * ts[hit]).astype(np.float32))}
    assert all(tmin_of[tuple(r)] == t for r, t in zip(out[redrawn, :3].view(np.uint32), out[redrawn, 3]))
    assert np.all(out[~empty, 7] == np.float32(3.4028235e38))
    d = out[~empty, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # the facing geometric normal at each origin: from its triangle (own hit rays)
    tri = v[idx[ids[own[same]]]].astype(np.float64)
    ng = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    flip = (ng * prim[own[same], 4:7]).sum(1) > 0
    ng[flip] *= -1
    c = (out[same, 4:7].astype(np.float64) * ng).sum(1)
    assert np.all(c > -1e-6)
    m = len(c)
    assert abs(c.mean() - 2 / 3) < 4 * np.sqrt(1 / 18 / m)       # var(cos) = 1/2 - 4/9
    assert abs((c * c).mean() - 0.5) < 4 * np.sqrt(1 / 12 / m)   # var(cos^2) = 1/3 - 1/4
    assert np.isfinite(origins[hit]).all()
