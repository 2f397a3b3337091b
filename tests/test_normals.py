"""Shading data of the scene (CPU): vertex normals and texture coordinates
through the scene API, the OBJ reader and the generators, and the oracle's
intersection records pinned by closed forms -- the interpolated shading
normal on an icosphere with analytic vertex normals is the sphere's normal
at the hit point, a mesh without normals shades with its geometric normal,
and the stored (u, v) rebuild the hit point."""
import numpy as np
import pytest

from conftest import random_rays

NO = 0xFFFFFFFF


def test_add_mesh_with_normals_and_texcoords(rbm):
    s = rbm.Scene()
    m = s.add_diffuse_bsdf((0.5, 0.5, 0.5))
    quad = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    qi = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    s.add_mesh(quad, qi, m)                                     # flat: no shading data
    n4 = np.array([[0, 0, 1], [0, 0.6, 0.8], [0.6, 0, 0.8], [0, 0, 1]], np.float32)
    s.add_mesh(quad, qi, m, normals=n4)                         # one normal per vertex
    uv = np.array([[0, 0], [1, 0], [1, 1]], np.float32)
    ni = np.array([[1, 0, 0], [NO, NO, NO]], np.uint32)         # second triangle: none
    ti = np.array([[0, 1, 2], [2, 1, 0]], np.uint32)
    s.add_mesh(quad, qi, m, normals=n4[:2], nidx=ni, texcoords=uv, tidx=ti)
    n, nidx, t, tidx = s.shading()
    assert n.shape == (6, 3) and t.shape == (3, 2)
    assert nidx.tolist() == [[NO] * 3, [NO] * 3, [0, 1, 2], [0, 2, 3], [5, 4, 4], [NO] * 3]
    assert tidx.tolist() == [[NO] * 3] * 4 + [[0, 1, 2], [2, 1, 0]]
    assert np.array_equal(n[:4], n4) and np.array_equal(n[4:], n4[:2]) and np.array_equal(t, uv)
    v, idx, _, _ = s.geometry()
    assert idx.tolist()[4:] == [[8, 9, 10], [8, 10, 11]]
    # refusals, nothing added
    nt = s.get_num_triangles()
    for bad in (dict(normals=n4[:3]),                                      # per vertex but 3 for 4 vertices
                dict(normals=n4, nidx=np.array([[0, 1, 4], [0, 1, 2]])),   # index past the normals
                dict(normals=n4, nidx=np.array([[0, NO, 2], [0, 1, 2]])),  # a partial triple
                dict(texcoords=uv)):                                       # per vertex but 3 for 4
        with pytest.raises(rbm.RbError):
            s.add_mesh(quad, qi, m, **bad)
        assert s.get_num_triangles() == nt
    # a scene without any shading data reports none
    s2 = rbm.Scene()
    s2.add_mesh(quad, qi, s2.add_diffuse_bsdf((0.5, 0.5, 0.5)))
    n, nidx, t, tidx = s2.shading()
    assert n.shape == (0, 3) and nidx is None and tidx is None


def test_obj_normals_and_texcoords(rbm, tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("\n".join([
        "v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0",
        "vt 0 0", "vt 1 0", "vt 1 1 0.5", "vt 0.25",
        "vn 0 0 1", "vn 0 0.6 0.8", "vn 0.6 0 0.8",
        "o a",
        "f 1/1/1 2/2/2 3/3/3",        # v/vt/vn
        "f 1//1 3//2 4//3",           # v//vn
        "f 1/1 2/2 3/3",              # v/vt
        "f 1 2 3",                    # v
        "o b",
        "f -4/-4/-3 -3/-3/-2 -2/-2/-1 -1/-1/-1",   # negative indices, a quad fanned in two
        "f 1/1/1 2/2 3/3/3",          # a corner without a normal: the triangle has none
    ]) + "\n")
    s = rbm.Scene()
    s.load_obj(p, s.add_diffuse_bsdf((0.5, 0.5, 0.5)))
    v, idx, _, _ = s.geometry()
    n, nidx, t, tidx = s.shading()
    assert idx.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 2], [0, 1, 2], [0, 1, 2], [0, 2, 3], [0, 1, 2]]
    assert nidx.tolist() == [[0, 1, 2], [0, 1, 2], [NO] * 3, [NO] * 3, [0, 1, 2], [0, 2, 2], [NO] * 3]
    assert tidx.tolist() == [[0, 1, 2], [NO] * 3, [0, 1, 2], [NO] * 3, [0, 1, 2], [0, 2, 3], [0, 1, 2]]
    assert np.allclose(n, [[0, 0, 1], [0, 0.6, 0.8], [0.6, 0, 0.8]])
    assert np.allclose(t, [[0, 0], [1, 0], [1, 1], [0.25, 0]])
    bad = tmp_path / "bad.obj"
    for face in ("f 1/1/9 2/2/2 3/3/3", "f 1/ 2/ 3/", "f 1/x/1 2 3"):
        bad.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvn 0 0 1\nvn 0 0 1\nvn 0 0 1\n" + face + "\n")
        with pytest.raises(rbm.RbError):
            _load(rbm, bad)


def _load(rbm, path):
    s = rbm.Scene()
    s.load_obj(path, s.add_diffuse_bsdf((0.5, 0.5, 0.5)))


def _smooth_spheres(rbm, count=6, seed=2):
    s = rbm.Scene()
    m = s.add_diffuse_bsdf((0.6, 0.6, 0.6))
    s.generate(rbm._lib.RB_GEN_ICOSPHERES, count=count, subdiv=3, seed=seed, r=(0.5, 1.2), c=(-3.0, 3.0),
               material_id=m, smooth=True)
    return s


def test_smooth_icospheres_carry_their_analytic_normals(rbm):
    s = _smooth_spheres(rbm)
    v, idx, _, _ = s.geometry()
    n, nidx, _, _ = s.shading()
    assert np.array_equal(nidx, idx)   # one normal per vertex, indexed like the vertices
    per = 642  # vertices of a subdivision-3 icosphere
    for k in range(len(v) // per):
        vs, ns = v[k * per:(k + 1) * per].astype(np.float64), n[k * per:(k + 1) * per].astype(np.float64)
        c = vs.mean(0)
        r = np.linalg.norm(vs - c, axis=1)
        assert np.allclose((vs - c) / r[:, None], ns, atol=2e-6)
        assert np.allclose(np.linalg.norm(ns, axis=1), 1.0, atol=1e-6)


def _trace_and_fill(rbm, orc, s, rays):
    from renderbox2_amd import _lib as L
    v, idx, tm, _ = s.geometry()
    b = rbm.BVH(s)
    nodes, woop, x = b.export(L.RB_LAYOUT_REF_COMPACT2)
    ids, ts = orc.trace_compact2(nodes.reshape(-1, 4), woop.reshape(-1, 4), x, rays)
    n, nidx, _, _ = s.shading()
    normals = orc.tri_normals(n, nidx, len(idx))
    return v, idx, ids, ts, orc.fill(rays, ids, ts, v, idx, tm, normals)


def test_oracle_fill_closed_forms(rbm, orc):
    """Flat mesh: Ns is Ng, bit for bit; Ng is the unit cross product;
    position o + t d; -d exact; epsilon 1e-3 t exact; front = Ng.d < 0; the
    stored (u, v) rebuild the position as u v0 + v v1 + (1 - u - v) v2."""
    s = rbm.scenes.ao_room(seed=3)
    rays = random_rays(20000, seed=8, box=4.0, tmax=3.0)
    v, idx, ids, ts, f = _trace_and_fill(rbm, orc, s, rays)
    h = ids >= 0
    assert 0.2 < h.mean() < 0.95
    tri = v[idx[ids[h]]].astype(np.float64)
    ng = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    assert np.allclose(f["geometric_normal"][h], ng, atol=1e-6)
    assert np.array_equal(f["shading_normal"][h].view(np.uint32), f["geometric_normal"][h].view(np.uint32))
    p = rays[h, :3].astype(np.float64) + rays[h, 4:7] * ts[h, None].astype(np.float64)
    assert np.allclose(f["position"][h], p, atol=1e-5)
    assert np.array_equal(f["incoming_direction"][h], -rays[h, 4:7])
    assert np.array_equal(f["epsilon"][h], (np.float32(1e-3) * ts[h]).astype(np.float32))
    assert np.array_equal(f["front"][h], ((ng * rays[h, 4:7]).sum(1) < 0).astype(np.uint32))
    u, w = f["uv"][h, 0].astype(np.float64), f["uv"][h, 1].astype(np.float64)
    rebuilt = u[:, None] * tri[:, 0] + w[:, None] * tri[:, 1] + (1 - u - w)[:, None] * tri[:, 2]
    assert np.allclose(rebuilt, f["position"][h], atol=2e-5)
    assert np.all((u > -1e-4) & (w > -1e-4) & (u + w < 1 + 1e-4))
    assert np.all(np.isnan(f["position"][~h]))  # misses untouched


def test_oracle_shading_normal_on_a_smooth_icosphere(rbm, orc):
    """On an icosphere whose vertex normals are the sphere's, the barycentric
    blend of the vertex normals is the direction of the hit point from the
    centre (u n0 + v n1 + w n2 = (p - c) / r when p = u v0 + v v1 + w v2), so
    Ns must equal (p - c) / |p - c| up to rounding; a wrong weight order (the
    reference's n0, n0, n1) or a missing normalisation fails it."""
    s = _smooth_spheres(rbm, count=1, seed=4)
    v, idx, _, _ = s.geometry()
    c = v.astype(np.float64).mean(0)
    rng = np.random.default_rng(1)
    o = c + rng.normal(0, 1, (5000, 3)) * 4.0
    d = c + rng.normal(0, 0.3, (5000, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((5000, 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = o, d, 3e38
    _, _, ids, ts, f = _trace_and_fill(rbm, orc, s, rays)
    h = ids >= 0
    assert h.mean() > 0.5
    p = f["position"][h].astype(np.float64)
    exact = (p - c) / np.linalg.norm(p - c, axis=1, keepdims=True)
    ns = f["shading_normal"][h].astype(np.float64)
    err = np.linalg.norm(ns - exact, axis=1)
    assert err.max() < 2e-5, err.max()
    assert np.allclose(np.linalg.norm(ns, axis=1), 1.0, atol=1e-6)
    # and it is not the flat normal (the faces are 0.1 rad apart)
    ng = f["geometric_normal"][h].astype(np.float64)
    assert np.median(np.linalg.norm(ns - ng, axis=1)) > 1e-3


def test_oracle_raycast_normals_image_is_the_sphere_normal(rbm, orc):
    """The ray-cast NORMALS mode shows the shading normal: on a smooth
    icosphere every covered pixel (1 sample, box film) is the unit direction
    of its hit point from the centre."""
    s = _smooth_spheres(rbm, count=1, seed=4)
    v, idx, tm, _ = s.geometry()
    n, nidx, _, _ = s.shading()
    from renderbox2_amd import _lib as L
    b = rbm.BVH(s)
    c2 = [a.reshape(-1, 4) if k < 2 else a for k, a in enumerate(b.export(L.RB_LAYOUT_REF_COMPACT2))]
    c = v.astype(np.float64).mean(0)
    cam = (tuple(c + [0.0, 0.0, -3.5]), tuple(c), (0.0, 1.0, 0.0), 40.0, 48, 48)
    normals = orc.tri_normals(n, nidx, len(idx))
    img = orc.render_raycast(c2, v, idx, tm, s.materials(), cam, (0, 0, 48, 48), 1, 3, 2, normals=normals)
    flat = orc.render_raycast(c2, v, idx, tm, s.materials(), cam, (0, 0, 48, 48), 1, 3, 2)
    cov = np.abs(img).sum(-1) > 0
    assert cov.mean() > 0.2
    lens = np.linalg.norm(img[cov].astype(np.float64), axis=1)
    assert np.allclose(lens, 1.0, atol=1e-6)
    assert np.abs(img - flat).max() > 1e-3     # smooth differs from flat
    # facing the camera: the normals point back towards it (z < 0 here)
    assert (img[cov][:, 2] < 0.05).all()
