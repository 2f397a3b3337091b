"""Host-side product code against the oracle (CPU only): the SBVH builder,
layouts, BVH text format, scene generators, camera chain, and the C-ABI
library's exported symbols."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, random_triangles

# ------------------------------------------------------------ C-ABI ---


def test_library_exports_every_declared_symbol(rbm):
    header = open(os.path.join(ROOT, "include", "rb.h")).read()
    names = set(re.findall(r"^\s*(?:rb_status|size_t|const char\*|void)\s+(rb_\w+)\s*\(", header, re.M))
    assert len(names) > 40
    lib = C.CDLL(os.path.join(ROOT, "renderbox2_amd", "librb.so"))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    from renderbox2_amd import _lib
    assert names <= set(_lib.exported_symbols())


def test_errors_are_reported_not_asserted(rbm):
    from renderbox2_amd import _lib as L
    lib = L.lib()
    assert lib.rb_scene_add_mesh(None, None, 0, None, 0, 0) == L.RB_ERR_INVALID
    assert "NULL" in L.last_error()
    s = rbm.Scene()
    m = s.add_diffuse_bsdf((0.5, 0.5, 0.5))
    with pytest.raises(rbm.RbError) as e:
        s.add_mesh(np.zeros((3, 3)), np.array([[0, 1, 5]]), m)
    assert e.value.status == L.RB_ERR_INVALID
    with pytest.raises(rbm.RbError):
        s.add_mesh(np.zeros((3, 3)), np.array([[0, 1, 2]]), 7)


# ------------------------------------------------------ SBVH builder ---


def _product_compact2(rbm, v, idx, max_leaf=None, threads=0, node_cost=1.0):
    from renderbox2_amd import _lib as L
    p = rbm.Platform(node_cost=node_cost)
    if max_leaf:
        p.set_leaf_preferences(1, max_leaf)
    b = rbm.BVH((v, idx), p, rbm.BuildParams(threads=threads))
    n, w, x = b.export(L.RB_LAYOUT_REF_COMPACT2)
    return b, n.reshape(-1, 4), w.reshape(-1, 4), x


@pytest.mark.parametrize("max_leaf", [None, 4])
def test_builder_matches_oracle_bytes(rbm, orc, max_leaf):
    v, idx = random_triangles(4000, seed=5)
    b, n, w, x = _product_compact2(rbm, v, idx, max_leaf)
    ob = orc.Bvh(v, idx, max_leaf=max_leaf or 0x7FFFFFF)
    on, ow, ox = ob.compact2()
    assert np.array_equal(n, on) and np.array_equal(w, ow) and np.array_equal(x, ox)
    ps, os_ = b.stats(), ob.stats()
    assert ps["sah"] == os_["sah"]
    assert (ps["n_inner"], ps["n_leaf"], ps["n_tri_refs"], ps["duplicates"]) == \
        (os_["n_inner"], os_["n_leaf"], os_["n_tris"], os_["duplicates"])


def test_builder_matches_oracle_on_generated_scene(rbm, orc):
    s = rbm.scenes.trace_scaled(12000, seed=9)
    v, idx, _, _ = s.geometry()
    _, n, w, x = _product_compact2(rbm, v, idx)
    on, ow, ox = orc.Bvh(v, idx).compact2()
    assert np.array_equal(n, on) and np.array_equal(w, ow) and np.array_equal(x, ox)


def test_builder_is_thread_count_independent(rbm):
    v, idx = random_triangles(30000, seed=6, spread=20.0)
    outs = [_product_compact2(rbm, v, idx, threads=t)[1:] for t in (1, 3, 8)]
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_parallel_sort_and_binning_match_the_serial_oracle(rbm, orc):
    """Nodes of >= 65,536 references sort and bin on several threads; the
    tree must still be the single-threaded reference's, byte for byte.
    (An empty per-thread spatial bin once leaked into the merged bins and
    changed split decisions at this size.)"""
    s = rbm.scenes.trace_scaled(80_000, seed=3)
    v, idx, _, _ = s.geometry()
    on, ow, ox = orc.Bvh(v, idx).compact2()
    for t in (2, 8):
        _, n, w, x = _product_compact2(rbm, v, idx, threads=t)
        assert np.array_equal(n, on) and np.array_equal(w, ow) and np.array_equal(x, ox), t


def test_builder_degenerates_and_small_scenes(rbm, orc):
    # one triangle: root leaf, wrapped by the exporters
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    from renderbox2_amd import _lib as L
    b = rbm.BVH((v, idx))
    n, w, x = b.export(L.RB_LAYOUT_MI355X)
    assert n.size == 16 and w.size == 12 and list(x) == [0]
    # lines are dropped like the reference does
    v2 = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 0, 0], [1, 0, 0], [0, 1, 0], [5, 5, 5], [6, 5, 5],
                   [5, 6, 5]], np.float32)
    idx2 = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 8]], np.uint32)
    _, n2, w2, x2 = _product_compact2(rbm, v2, idx2, node_cost=0.1)
    on, ow, ox = orc.Bvh(v2, idx2, node_cost=0.1).compact2()
    assert len(np.unique(ox[np.arange(len(ox)) % 3 == 0])) >= 2
    assert np.array_equal(n2, on) and np.array_equal(w2, ow) and np.array_equal(x2, ox)


# ------------------------------------------------------------ layouts ---


def _decode_mi355x(nodes, tris, index):
    nodes = nodes.reshape(-1, 16)
    leaves = []
    stack = [0]
    seen = 0
    while stack:
        r = stack.pop()
        seen += 1
        rec = nodes[r]
        for k in range(2):
            link = np.int32(rec[12 + k])
            if link >= 0:
                stack.append(int(link))
            else:
                l = ~int(link) & 0xFFFFFFFF
                first, cnt = l >> 5, l & 31
                assert cnt == rec[14 + k]
                leaves.append(sorted(index[first:first + cnt].tolist()))
    return leaves, seen


def test_mi355x_layout_is_the_same_tree(rbm, orc):
    from renderbox2_amd import _lib as L
    v, idx = random_triangles(3000, seed=12)
    b = rbm.BVH((v, idx))
    n, t, i = b.export(L.RB_LAYOUT_MI355X)
    leaves, nrec = _decode_mi355x(n, t, i)
    cn, cw, cx = b.export(L.RB_LAYOUT_REF_COMPACT2)
    from test_oracle import _decode_compact2
    cl = _decode_compact2(cn.reshape(-1, 4), cw.reshape(-1, 4), cx)
    assert sorted(leaves) == sorted(sorted(ids) for ids, _ in cl)
    assert nrec == cn.size // 16
    # Woop rows equal the Compact2 rows (minus terminators)
    cw4 = cw.reshape(-1, 4)
    keep = np.ones(len(cw4), bool)
    term = np.where(cw4[:, 0] == 0x80000000)[0]
    keep[term] = False
    assert np.array_equal(np.sort(t.reshape(-1, 12).view(np.uint32), axis=0),
                          np.sort(cw4[keep].reshape(-1, 12), axis=0))


def test_mi355x_splits_big_leaves(rbm):
    from renderbox2_amd import _lib as L
    # 100 identical triangles cannot be separated: one leaf of 100 refs,
    # exported as sub-leaves of <= 31.
    v = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (100, 1))
    idx = np.arange(300, dtype=np.uint32).reshape(100, 3)
    b = rbm.BVH((v, idx))
    n, t, i = b.export(L.RB_LAYOUT_MI355X)
    leaves, _ = _decode_mi355x(n, t, i)
    assert all(len(l) <= 31 for l in leaves)
    refs = sum(leaves, [])
    assert sorted(set(refs)) == list(range(100))
    assert len(refs) == b.stats()["n_tri_refs"] == len(i)
    # a stack of identical triangles ends in at least one leaf the exporter had to split
    cn, cw, cx = b.export(L.RB_LAYOUT_REF_COMPACT2)
    from test_oracle import _decode_compact2
    assert max(len(ids) for ids, _ in _decode_compact2(cn.reshape(-1, 4), cw.reshape(-1, 4), cx)) > 31


def test_bvh_text_format_round_trip(rbm, tmp_path):
    from renderbox2_amd import _lib as L
    v, idx = random_triangles(500, seed=13)
    b = rbm.BVH((v, idx))
    p = str(tmp_path / "scene.bvh")
    b.save_text(p)
    n, w, x = rbm.CudaBvh.load_text(p)
    cb = rbm.CudaBvh(b, L.RB_LAYOUT_REF_COMPACT2)
    assert np.array_equal(n, cb.get_node_buffer()) and np.array_equal(w, cb.get_tri_woop_buffer())
    assert np.array_equal(x, cb.get_tri_index_buffer())
    head = open(p).read().split("\n")[:5]
    assert head[0] == str(0x76543210) and head[4] == "1"
    # a truncated file is rejected
    txt = open(p).read()
    open(p, "w").write(txt[: len(txt) // 2])
    with pytest.raises(rbm.RbError):
        rbm.CudaBvh.load_text(p)


# --------------------------------------------------------- generators ---


def test_generated_scene_sizes_are_exact(rbm):
    assert rbm.scenes.ao_room().get_num_triangles() == 78_848
    assert rbm.scenes.trace_scaled(10_000).get_num_triangles() == 10_000
    assert rbm.scenes.trace_scaled(200_000, seed=2).get_num_triangles() == 200_000
    s = rbm.scenes.pt_room()
    nv, nt, nl, nm = s.counts()
    assert nt == 2_000_136 and nl == 8
    kinds = [m[0] for m in s.materials()]
    assert kinds.count(2) == 1 and kinds.count(1) > 200 and kinds.count(0) > 600


def test_room_faces_inward_and_lights_face_down(rbm):
    s = rbm.scenes.pt_room(spheres_hi=0, spheres_lo=1)
    v, idx, tm, lt = s.geometry()
    tri = v[idx]
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert len(lt) == 8 and np.allclose(n[lt], [0, -1, 0], atol=1e-6)
    area = 0.5 * np.linalg.norm(np.cross(tri[lt, 1] - tri[lt, 0], tri[lt, 2] - tri[lt, 0]), axis=1)
    assert np.allclose(area.reshape(4, 2).sum(1), 0.25, rtol=1e-5)
    room = np.arange(2048)
    # inward: the normal points towards the room centre
    assert np.all((n[room] * (0 - tri[room].mean(1))).sum(1) > 0)


def test_generators_are_deterministic(rbm):
    a = rbm.scenes.trace_scaled(20_000, seed=4).geometry()
    b = rbm.scenes.trace_scaled(20_000, seed=4).geometry()
    c = rbm.scenes.trace_scaled(20_000, seed=5).geometry()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])


def test_icosphere_geometry(rbm):
    from renderbox2_amd import _lib as L
    s = rbm.Scene()
    m = s.add_diffuse_bsdf((1, 1, 1))
    s.generate(L.RB_GEN_ICOSPHERES, count=1, subdiv=3, seed=1, r=(2.0, 2.0), c=(0.0, 0.0), material_id=m)
    v, idx, _, _ = s.geometry()
    assert len(idx) == 1280 and len(v) == 642
    assert np.allclose(np.linalg.norm(v, axis=1), 2.0, atol=1e-5)
    # outward winding: normals point away from the centre
    tri = v[idx]
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    assert np.all((n * tri.mean(1)).sum(1) > 0)


# -------------------------------------------------------------- camera ---


def test_camera_matrices_match_oracle(rbm, orc):
    lib = orc.lib()
    P = C.c_void_p
    lib.orc_camera_matrices.argtypes = [P, P, P, C.c_float, C.c_uint32, C.c_uint32, P, P]
    for pos, at, fov, w, h in [((0, 0.5, -4.8), (0, 0, 0), 70.0, 512, 512),
                               ((0, 4, -28), (0, 0, 0), 50.0, 2048, 1024),
                               ((3, 2, 1), (-1, 0.5, 2), 33.0, 640, 480)]:
        cam = rbm.PerspectiveCamera(pos, at, (0, 1, 0), (w, h), fov)
        r2c, c2w = cam.matrices()
        a = np.empty(16, np.float32)
        b = np.empty(16, np.float32)
        p3 = [np.array(x, np.float32) for x in (pos, at, (0, 1, 0))]
        lib.orc_camera_matrices(*[x.ctypes.data_as(P) for x in p3], fov, w, h, a.ctypes.data_as(P),
                                b.ctypes.data_as(P))
        assert np.array_equal(r2c.ravel(), a) and np.array_equal(c2w.ravel(), b)


def test_camera_geometry_closed_form(orc):
    # centre pixel looks along (lookat - position); horizontal fov spans the width
    lib = orc.lib()
    P = C.c_void_p
    lib.orc_camera_matrices.argtypes = [P, P, P, C.c_float, C.c_uint32, C.c_uint32, P, P]
    lib.orc_raygen_camera.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint64, C.c_int, P]
    pos, at, up = (np.array(x, np.float32) for x in ((1, 2, 3), (1, 2, 13), (0, 1, 0)))
    r2c = np.empty(16, np.float32)
    c2w = np.empty(16, np.float32)
    W, H = 200, 100
    lib.orc_camera_matrices(pos.ctypes.data_as(P), at.ctypes.data_as(P), up.ctypes.data_as(P), 90.0, W, H,
                            r2c.ctypes.data_as(P), c2w.ctypes.data_as(P))
    rays = np.empty((W * H, 8), np.float32)
    lib.orc_raygen_camera(r2c.ctypes.data_as(P), c2w.ctypes.data_as(P), W, 0, 0, W, H, 1, 0, 0,
                          rays.ctypes.data_as(P))
    rays = rays.reshape(H, W, 8)
    assert np.allclose(rays[0, 0, :3], pos)
    d = rays[:, :, 4:7]
    assert np.allclose(np.linalg.norm(d, axis=2), 1, atol=1e-6)
    mid = (d[H // 2 - 1, W // 2 - 1] + d[H // 2, W // 2]) / 2
    assert np.allclose(mid / np.linalg.norm(mid), [0, 0, 1], atol=1e-3)
    # 90 degrees: the film spans [-1, 1] at unit distance, so the centre of
    # the first / last pixel column sits at tan(angle) = 1 - 1/W
    left = d[H // 2, 0]
    right = d[H // 2, W - 1]
    ang = np.degrees(np.arccos(np.clip([left[2], right[2]], -1, 1)))
    assert np.allclose(ang, np.degrees(np.arctan(1 - 1.0 / W)), atol=0.02)
    # the reference maps y onto [-1, 1] as well (both axes use the same scale)
    top = d[0, W // 2 - 1] + d[0, W // 2]
    assert abs(np.degrees(np.arctan2(top[1], top[2])) - np.degrees(np.arctan(1 - 1.0 / H))) < 0.05
    # image y grows downward, x mirrors so that +x in raster is to the camera's right
    assert d[0, W // 2, 1] > 0 > d[H - 1, W // 2, 1]


# ------------------------------------------------------- binary BVH cache ---


def test_binary_bvh_cache_round_trip_and_key_checks(rbm, tmp_path):
    from renderbox2_amd import _lib as L
    s = rbm.scenes.pt_scaled(30_000, seed=2)
    b = rbm.BVH(s)
    path = tmp_path / "room.rbvh"
    b.save_binary(path)
    c = rbm.BVH.load_binary(path, s)
    for layout in (L.RB_LAYOUT_REF_COMPACT2, L.RB_LAYOUT_MI355X):
        for x, y in zip(b.export(layout), c.export(layout)):
            assert np.array_equal(x, y)
    assert b.stats() == c.stats()
    # other build parameters: refused
    p = rbm.Platform()
    p.set_leaf_preferences(1, 4)
    with pytest.raises(rbm.RbError):
        rbm.BVH.load_binary(path, s, p)
    # other geometry: refused
    s2 = rbm.scenes.pt_scaled(30_000, seed=3)
    with pytest.raises(rbm.RbError):
        rbm.BVH.load_binary(path, s2)
    # corrupt payload (one byte) and truncation: refused
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 0x40
    bad = tmp_path / "bad.rbvh"
    bad.write_bytes(bytes(raw))
    with pytest.raises(rbm.RbError):
        rbm.BVH.load_binary(bad, s)
    bad.write_bytes(path.read_bytes()[:-100])
    with pytest.raises(rbm.RbError):
        rbm.BVH.load_binary(bad, s)
    # a header whose node count does not match the file (here 2^31 nodes,
    # 64 GiB if allocated) is refused before anything is allocated; the
    # count sits after magic, version, reserved, key, nv and nt
    raw = bytearray(path.read_bytes())
    raw[32:36] = (1 << 31).to_bytes(4, "little")
    bad.write_bytes(bytes(raw))
    with pytest.raises(rbm.RbError, match="header describes"):
        rbm.BVH.load_binary(bad, s)


def test_build_cached_builds_once_then_loads(rbm, tmp_path):
    from renderbox2_amd import _lib as L
    s = rbm.scenes.trace_scaled(20_000, seed=4)
    b1, hit1 = rbm.BVH.build_cached(s, tmp_path)
    b2, hit2 = rbm.BVH.build_cached(s, tmp_path)
    assert (hit1, hit2) == (False, True)
    assert len(list(tmp_path.glob("rbvh_*.bin"))) == 1
    assert not list(tmp_path.glob("*.tmp"))
    for x, y in zip(b1.export(L.RB_LAYOUT_MI355X), b2.export(L.RB_LAYOUT_MI355X)):
        assert np.array_equal(x, y)
    # the thread count does not enter the key (the tree does not depend on it)
    _, hit3 = rbm.BVH.build_cached(s, tmp_path, params=rbm.BuildParams(threads=3))
    assert hit3


def test_product_builder_matches_the_golden_fixture(rbm):
    """librb.so's parallel builder and Compact2 exporter give the bytes the
    committed fixture holds for the seeded 2k-triangle scene."""
    import hashlib
    import importlib.util
    import os
    from renderbox2_amd import _lib as L
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    spec = importlib.util.spec_from_file_location("make_fixtures", os.path.join(golden, "make_fixtures.py"))
    mf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mf)
    g = np.load(os.path.join(golden, "trace_2k.npz"))
    v, idx, _ = mf.trace_2k_inputs()
    p = rbm.Platform()
    p.set_leaf_preferences(1, mf.SCENE["max_leaf"])
    for threads in (1, 8):
        b = rbm.BVH((v, idx), p, rbm.BuildParams(threads=threads))
        n, w, x = b.export(L.RB_LAYOUT_REF_COMPACT2)
        assert [n.size, w.size, x.size] == g["sizes"].tolist()
        for name, a in (("sha_nodes", n), ("sha_woop", w), ("sha_index", x)):
            assert hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest() == g[name].tobytes(), (name, threads)


# ------------------------------------------- refusals before any read ---


def test_scene_growth_limit_is_checked_before_the_arrays(rbm):
    """2^31 triangles are refused from the counts alone: the arrays (3
    vertices, 1 triangle here) are never read past, by add_mesh and by the
    OBJ reader alike (both take the same limit check)."""
    from renderbox2_amd import _lib as L
    lib = L.lib()
    s = rbm.Scene()
    m = s.add_diffuse_bsdf((0.5, 0.5, 0.5))
    v = np.zeros((3, 3), np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    st = lib.rb_scene_add_mesh(s._h, v.ctypes.data, 3, idx.ctypes.data, 1 << 31, m)
    assert st == L.RB_ERR_INVALID and "2^31" in L.last_error()
    assert s.get_num_triangles() == 0
    assert lib.rb_scene_add_mesh(s._h, v.ctypes.data, 3, idx.ctypes.data, 1, m) == L.RB_OK
    assert s.get_num_triangles() == 1


class _Hash64:
    """The binary cache's checksum (rb_bvh_cache.cpp), for writing test files."""
    M = (1 << 64) - 1

    def __init__(self):
        self.h, self.n = 0x9E3779B97F4A7C15, 0

    @staticmethod
    def _rotl(x, r):
        return ((x << r) | (x >> (64 - r))) & _Hash64.M

    def bytes(self, b):
        b = bytes(b)
        for i in range(0, len(b), 8):
            w = int.from_bytes(b[i:i + 8].ljust(8, b"\0"), "little")
            self.h ^= (self._rotl(w * 0xC2B2AE3D27D4EB4F & self.M, 31) * 0x9E3779B185EBCA87) & self.M
            self.h = (self._rotl(self.h, 27) * 0x9E3779B185EBCA87 + 0x85EBCA77C2B2AE63) & self.M
        self.n += len(b)

    def done(self):
        x = self.h ^ self.n
        x ^= x >> 33
        x = x * 0xFF51AFD7ED558CCD & self.M
        x ^= x >> 33
        x = x * 0xC4CEB9FE1A85EC53 & self.M
        return x ^ (x >> 33)


def _cache_file_with_nodes(good_bytes, nodes, refs):
    """A cache file with the header (key, sizes, stats) of a real one and
    the given node records / reference list, with a valid checksum."""
    import struct
    hdr = bytearray(good_bytes[:100])  # the packed 100-byte header
    struct.pack_into("<II", hdr, 32, len(nodes), len(refs))
    body = b"".join(struct.pack("<6fII", *lo, *hi, lc, link) for lo, hi, lc, link in nodes)
    body += np.asarray(refs, np.uint32).tobytes()
    h = _Hash64()
    h.bytes(hdr)
    h.bytes(body[:len(nodes) * 32])
    h.bytes(body[len(nodes) * 32:])
    return bytes(hdr) + body + h.done().to_bytes(8, "little")


def test_binary_cache_refuses_files_no_build_writes(rbm, tmp_path):
    """The checksum only detects corruption; a well-formed file must also be
    a tree (every node but the root linked once) no deeper than the builder
    goes (64), or the traversal stack could overflow."""
    s = rbm.scenes.trace_scaled(2_000, seed=1)
    b = rbm.BVH(s)
    path = tmp_path / "good.rbvh"
    b.save_binary(path)
    good = path.read_bytes()
    assert len(good) > 100 and good[:8] == b"RBVHBIN1"
    box = ((-20.0,) * 3, (20.0,) * 3)
    leaf = lambda first: (*box, 0x80000001, first)  # noqa: E731

    def chain(depth):
        """Inner nodes 0..depth-1, each with the next as first child and a
        one-reference leaf as second child (preorder, as the writer emits)."""
        nodes, refs = [], []

        def emit(d):
            me = len(nodes)
            if d == depth:
                nodes.append(leaf(len(refs)))
                refs.append(0)
                return
            nodes.append([*box, 0, 0])
            emit(d + 1)
            nodes[me][3] = len(nodes)
            nodes.append(leaf(len(refs)))
            refs.append(1)

        emit(0)
        return [tuple(n) for n in nodes], refs

    # the same file rewritten by this writer loads (the writer is right) ...
    ok_nodes, ok_refs = chain(64)
    p = tmp_path / "deep64.rbvh"
    p.write_bytes(_cache_file_with_nodes(good, ok_nodes, ok_refs))
    rbm.BVH.load_binary(p, s)
    # ... one level deeper does not
    deep_nodes, deep_refs = chain(65)
    p.write_bytes(_cache_file_with_nodes(good, deep_nodes, deep_refs))
    with pytest.raises(rbm.RbError, match="deeper than a build makes"):
        rbm.BVH.load_binary(p, s)
    # a node linked twice (a DAG, not a tree)
    dag = [(*box, 0, 3), (*box, 0, 3), leaf(0), leaf(0)]
    p.write_bytes(_cache_file_with_nodes(good, dag, [0]))
    with pytest.raises(rbm.RbError, match="not a tree"):
        rbm.BVH.load_binary(p, s)


def test_product_builds_the_single_thread_oracle_tree_at_one_million_triangles(rbm):
    """The headline scene's SBVH (1,000,000 triangles, built on every CPU)
    has the bytes of the single-threaded oracle's tree, whose SHA-256 the
    fixture holds (tests/golden/make_fixtures.py): the tree behind the
    benchmark's visit counts and algorithmic bytes is the reference's."""
    import hashlib
    import json
    from renderbox2_amd import _lib as L
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "trace_1m_tree.json")))
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    s = rbm.scenes.trace_1m(seed=0)
    v, idx, _, _ = s.geometry()
    assert (sha(v), sha(idx)) == (g["sha_verts"], g["sha_idx"]), "the scene generator changed"
    b = rbm.BVH(s)
    n, w, x = b.export(L.RB_LAYOUT_REF_COMPACT2)
    assert [n.size, w.size, x.size] == g["sizes"]
    assert (sha(n), sha(w), sha(x)) == (g["sha_nodes"], g["sha_woop"], g["sha_index"])
    st = b.stats()
    assert (st["n_inner"], st["n_leaf"], st["n_tri_refs"], st["duplicates"]) == \
        (g["stats"]["n_inner"], g["stats"]["n_leaf"], g["stats"]["n_tris"], g["stats"]["duplicates"])
