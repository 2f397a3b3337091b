"""Closed-form pins of the oracle's AO and path-tracing passes (CPU)."""
import ctypes as C

import numpy as np
import pytest


def _quad(x0, x1, z0, z1, y, up=True):
    v = np.array([[x0, y, z0], [x1, y, z0], [x1, y, z1], [x0, y, z1]], np.float32)
    # (v1-v0) x (v2-v0) = +y for this winding when up
    idx = np.array([[0, 2, 1], [0, 3, 2]] if up else [[0, 1, 2], [0, 2, 3]], np.uint32)
    return v, idx


def _merge(*parts):
    vs, ids, off = [], [], 0
    for v, i in parts:
        vs.append(v)
        ids.append(i + off)
        off += len(v)
    return np.concatenate(vs), np.concatenate(ids)


def _primary(orc, c2, cam, spp, seed):
    r2c, c2w = orc.camera_matrices(*cam)
    w, h = cam[4], cam[5]
    lib = orc.lib()
    P = C.c_void_p
    lib.orc_raygen_camera.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint64, C.c_int, P]
    rays = np.empty((w * h * spp, 8), np.float32)
    lib.orc_raygen_camera(r2c.ctypes.data_as(P), c2w.ctypes.data_as(P), w, 0, 0, w, h, spp, seed, 1,
                          rays.ctypes.data_as(P))
    ids, t = orc.trace_compact2(*c2, rays)
    return rays, ids, t


def test_ao_open_plane_is_one(orc):
    v, idx = _quad(-50, 50, -50, 50, 0.0)
    b = orc.Bvh(v, idx)
    cam = ((0.0, 3.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 48, 32)
    img = orc.render_ao(b.compact2(), v, idx, cam, (0, 0, 48, 32), 1, 3, 8, 1.0)
    vals = np.unique(img)
    assert set(vals.tolist()) <= {0.0, 1.0}
    assert (img == 1.0).mean() > 0.5


def test_ao_inside_closed_sphere_is_zero(orc, rbm):
    from renderbox2_amd import _lib as L
    s = rbm.Scene()
    m = s.add_diffuse_bsdf((1, 1, 1))
    s.generate(L.RB_GEN_ICOSPHERES, count=1, subdiv=2, seed=0, r=(0.5, 0.5), c=(0.0, 0.0), material_id=m)
    v, idx, _, _ = s.geometry()
    b = orc.Bvh(v, idx)
    cam = ((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), 90.0, 32, 32)
    img = orc.render_ao(b.compact2(), v, idx, cam, (0, 0, 32, 32), 2, 5, 8, 1.5)
    assert np.all(img == 0.0)


def test_ao_floor_wall_wedge_closed_form(orc):
    # floor y = 0 for x < 0, wall x = 0 facing -x; AO radius r.  From a floor
    # point at distance d from the wall, cosine-weighted rays of length r
    # reach the wall iff their x component exceeds a = d / r; for a uniform
    # disk point that has probability (acos(a) - a sqrt(1 - a^2)) / pi.
    floor = _quad(-50, 0, -50, 50, 0.0)
    wv = np.array([[0, 0, -50], [0, 0, 50], [0, 50, 50], [0, 50, -50]], np.float32)
    wall = (wv, np.array([[0, 1, 2], [0, 2, 3]], np.uint32))
    v, idx = _merge(floor, wall)
    b = orc.Bvh(v, idx)
    c2 = b.compact2()
    r = 1.0
    cam = ((-0.75, 4.0, 0.0), (-0.75, 0.0, 0.0), (0.0, 0.0, 1.0), 30.0, 64, 64)
    spp, nao, seed = 4, 16, 9
    img = orc.render_ao(c2, v, idx, cam, (0, 0, 64, 64), spp, seed, nao, r)
    rays, ids, t = _primary(orc, c2, cam, spp, seed)
    hitp = rays[:, :3] + rays[:, 4:7] * t[:, None]
    on_floor = (ids >= 0) & (ids < 2) & (np.abs(hitp[:, 1]) < 1e-4)
    a = np.clip(-hitp[:, 0] / r, 0.0, 1.0)
    expect = 1.0 - (np.arccos(a) - a * np.sqrt(1 - a * a)) / np.pi
    # pixels whose spp samples all land on the floor
    pix_floor = on_floor.reshape(-1, spp).all(1)
    exp_pix = expect.reshape(-1, spp).mean(1)
    var_pix = (expect * (1 - expect) / nao).reshape(-1, spp).sum(1) / spp ** 2
    got = img.reshape(-1, 3)[:, 0]
    sel = pix_floor
    assert sel.sum() > 1500
    diff = got[sel] - exp_pix[sel]
    sigma = np.sqrt(var_pix[sel].sum()) / sel.sum()
    assert abs(diff.mean()) < 4 * sigma + 1e-4
    # and by distance bins
    dsel = -hitp[:, 0].reshape(-1, spp).mean(1)[sel]
    for lo, hi in [(0, 0.3), (0.3, 0.6), (0.6, 1.2)]:
        k = (dsel >= lo) & (dsel < hi)
        if k.sum() > 50:
            s = np.sqrt(var_pix[sel][k].sum()) / k.sum()
            assert abs(diff[k].mean()) < 4 * s + 1e-4


def _direct_light_closed_form(p, rho, le, lv0, lv1, lv2, lv3, n=64):
    # (rho / pi) * Le * integral over the light quad of cos_p cos_l / r^2 dA
    uu = (np.arange(n) + 0.5) / n
    a = lv1 - lv0
    c = lv3 - lv0
    area = np.linalg.norm(np.cross(a, c))
    ln = np.cross(a, c) / area
    pts = lv0 + uu[:, None, None] * a + uu[None, :, None] * c
    d = pts - p
    r2 = (d * d).sum(-1)
    dn = d / np.sqrt(r2)[..., None]
    cos_p = np.clip(dn[..., 1], 0, None)
    cos_l = np.clip(-(dn * ln).sum(-1), 0, None)
    return rho / np.pi * le * (cos_p * cos_l / r2).mean() * area


def test_pt_direct_light_closed_form(orc):
    rho, le = 0.5, 10.0
    floor = _quad(-20, 20, -20, 20, 0.0)
    L = 0.5
    lq = np.array([[-L, 2, -L], [L, 2, -L], [L, 2, L], [-L, 2, L]], np.float32)
    light = (lq, np.array([[0, 1, 2], [0, 2, 3]], np.uint32))  # (v1-v0)x(v2-v0) = -y: faces down
    v, idx = _merge(floor, light)
    tri_mat = np.array([0, 0, 1, 1], np.uint32)
    mats = [(0, (rho, rho, rho)), (2, (le, le, le))]
    lights = np.array([2, 3], np.uint32)
    b = orc.Bvh(v, idx)
    c2 = b.compact2()
    cam = ((0.0, 1.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 32, 32)
    spp, seed = 16, 2
    img = orc.render_pt(c2, v, idx, tri_mat, mats, lights, cam, (0, 0, 32, 32), spp, seed, 0, 3, 0.5)
    rays, ids, t = _primary(orc, c2, cam, spp, seed)
    hitp = rays[:, :3] + rays[:, 4:7] * t[:, None]
    floor_hit = ((ids == 0) | (ids == 1)).reshape(-1, spp).all(1)
    exp = np.array([_direct_light_closed_form(hp.astype(np.float64), rho, le, *lq.astype(np.float64))
                    for hp in hitp])
    exp_pix = exp.reshape(-1, spp).mean(1)
    got = img.reshape(-1, 3)[:, 0]
    sel = floor_hit & (exp_pix > 0.05)
    assert sel.sum() > 300
    rel = (got[sel] - exp_pix[sel]) / exp_pix[sel]
    assert abs(rel.mean()) < 0.02
    assert np.percentile(np.abs(rel), 90) < 0.25
    assert np.allclose(img[..., 0], img[..., 1]) and np.allclose(img[..., 0], img[..., 2])


def test_pt_sees_emitter_and_its_mirror_image(orc):
    le, refl = 7.0, 0.8
    mirror = _quad(-20, 20, -20, 20, 0.0)
    lq = np.array([[-1, 2, -1], [1, 2, -1], [1, 2, 1], [-1, 2, 1]], np.float32)
    light = (lq, np.array([[0, 1, 2], [0, 2, 3]], np.uint32))
    v, idx = _merge(mirror, light)
    tri_mat = np.array([0, 0, 1, 1], np.uint32)
    mats = [(1, (refl, refl, refl)), (2, (le, le, le))]
    b = orc.Bvh(v, idx)
    # camera between the mirror (y = 0) and the 2x2 light (y = 2) looking
    # down: the light's image sits 3 below the camera, half-angle atan(1/3)
    # = 18.4 deg, so a 20 deg field of view sees only that image
    cam = ((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 20.0, 16, 16)
    img = orc.render_pt(b.compact2(), v, idx, tri_mat, mats, np.array([2, 3], np.uint32), cam, (0, 0, 16, 16), 1,
                        0, 3, 3, 0.5)
    # every pixel: mirror at bounce 0 (no direct light: specular), reflected
    # ray hits the light from below -> refl * Le (through the specular chain)
    assert np.allclose(img, refl * le, rtol=1e-5)
    cam2 = ((0.0, 1.0, 0.0), (0.0, 3.0, 0.0), (0.0, 0.0, 1.0), 40.0, 8, 8)
    img2 = orc.render_pt(b.compact2(), v, idx, tri_mat, mats, np.array([2, 3], np.uint32), cam2, (0, 0, 8, 8), 1,
                         0, 3, 3, 0.5)
    assert np.all(img2 == le)  # the light's emitting face seen directly


# ---------------------------------------------------------------- BDPT ----

@pytest.mark.parametrize("mask,n,max_depth,lvc_depth,expect", [
    (0b0, 1, 8, 4, 1),    # the camera sees the light: emitter hit only
    (0b0, 2, 8, 4, 1),    # one diffuse vertex: direct light only, no interior edge
    (0b0, 3, 8, 4, 2),    # direct + edge (z1, z2)
    (0b0, 5, 8, 4, 4),    # direct + edges 1, 2, 3 (light ends at depths 2, 1, 0)
    (0b0, 5, 8, 2, 3),    # edge 1 would need a light vertex of depth 2: not cached
    (0b1, 4, 8, 4, 2),    # z1 specular: edge 1 gone, edge 2 and direct remain
    (0b10, 4, 8, 4, 1),   # z2 specular: edges 1 and 2 both touch it
    (0b100, 4, 8, 4, 2),  # z3 specular: emitter hit after the mirror + edge 1
    (0b100, 4, 2, 4, 1),  # ... but that emitter hit is bounce 3 > max_depth 2
    (0b0, 6, 0, 8, 1),    # only one camera vertex: edge 1 is the only technique
    (0b0, 6, 0, 3, 0),    # ... and it needs a depth-3 light vertex: unbuildable
])
def test_bdpt_technique_counts_by_hand(orc, mask, n, max_depth, lvc_depth, expect):
    assert orc.bdpt_techniques(mask, n, max_depth, lvc_depth) == expect


def _box_room(mirror_wall=True):
    def q(p0, p1, p2, p3):
        return np.array([p0, p1, p2, p3], np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    walls = [q([-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]),    # floor
             q([-1, 2, -1], [1, 2, -1], [1, 2, 1], [-1, 2, 1]),    # ceiling
             q([-1, 0, 1], [1, 0, 1], [1, 2, 1], [-1, 2, 1]),      # back
             q([-1, 0, -1], [1, 0, -1], [1, 2, -1], [-1, 2, -1]),  # front
             q([-1, 0, -1], [-1, 0, 1], [-1, 2, 1], [-1, 2, -1]),  # left
             q([1, 0, -1], [1, 0, 1], [1, 2, 1], [1, 2, -1])]      # right
    a = 0.3
    light = q([-a, 1.98, -a], [a, 1.98, -a], [a, 1.98, a], [-a, 1.98, a])  # normal -y: faces the floor
    v, idx = _merge(*walls, light)
    right = 1 if mirror_wall else 0
    tri_mat = np.array([0] * 10 + [right] * 2 + [2] * 2, np.uint32)
    mats = [(0, (0.7, 0.6, 0.5)), (1, (0.9, 0.9, 0.9)), (2, (8.0, 8.0, 8.0))]
    return v, idx, tri_mat, mats, np.array([12, 13], np.uint32)


def test_bdpt_without_connections_is_the_path_tracer(orc):
    # no connections -> every technique weight is 1 -> the camera pass is the
    # path tracer, bit for bit, whatever the light pass stored
    v, idx, tm, mats, lights = _box_room()
    c2 = orc.Bvh(v, idx).compact2()
    cam = ((0.0, 1.0, -0.95), (0.0, 1.0, 1.0), (0.0, 1.0, 0.0), 70.0, 16, 16)
    pt = orc.render_pt(c2, v, idx, tm, mats, lights, cam, (0, 0, 16, 16), 4, 3, 8, 3, 0.5)
    bd = orc.render_bdpt(c2, v, idx, tm, mats, lights, cam, (0, 0, 16, 16), 4, 3, 8, 3, 0.5, 256, 4, 0)
    assert np.array_equal(pt.view(np.uint32), bd.view(np.uint32))


def test_bdpt_agrees_with_the_path_tracer_in_expectation(orc):
    # Both estimators are unbiased for the same image once the depth limits
    # no longer truncate (albedo <= 0.9, roulette from bounce 3: paths longer
    # than 26 vertices carry < 1e-4 of the energy).  A closed box with a
    # mirror wall exercises the specular masks.  8 seeds x 512 spp put the
    # channel means' standard errors near 0.25%; mistakes in the weights or
    # the connection estimator move them by 5-20%.
    v, idx, tm, mats, lights = _box_room()
    c2 = orc.Bvh(v, idx).compact2()
    W = H = 24
    cam = ((0.0, 1.0, -0.95), (0.0, 1.0, 1.0), (0.0, 1.0, 0.0), 70.0, W, H)
    pt, bd = [], []
    for seed in range(8):
        pt.append(orc.render_pt(c2, v, idx, tm, mats, lights, cam, (0, 0, W, H), 512, seed, 24, 3, 0.5)
                  .reshape(-1, 3).mean(0))
        bd.append(orc.render_bdpt(c2, v, idx, tm, mats, lights, cam, (0, 0, W, H), 512, seed, 24, 3, 0.5, 4096, 4, 3)
                  .reshape(-1, 3).mean(0))
    pt, bd = np.mean(pt, 0), np.mean(bd, 0)
    assert np.all(pt > 0.1)
    assert np.all(np.abs(bd / pt - 1.0) < 0.01), (pt, bd)


# ---------------------------------------------------------------- film ----

def _jitter(orc, seed, width, gx, gy, s):
    key = [seed & 0xFFFFFFFF, ((seed >> 32) ^ 0x85EBCA6B) & 0xFFFFFFFF]  # stream tag 1 (camera)
    q = orc.philox([gy * width + gx, s, 0, 0], key)
    return (q[0] >> 8) * 2.0 ** -24, (q[1] >> 8) * 2.0 ** -24


def _filter64(kind, xw, yw, alpha, x, y):
    if kind == 1:
        return max(0.0, xw - abs(x)) * max(0.0, yw - abs(y))
    if kind == 2:
        g = lambda d, w: max(0.0, np.exp(-alpha * d * d) - np.exp(-alpha * w * w))  # noqa: E731
        return g(x, xw) * g(y, yw)
    return 1.0


def _splat64(orc, samples, width, tile, spp, seed, film):
    """kernel_update_film's splat (samplerrenderer.cu:103-169) in float64:
    every sample adds w*L and w to the tile pixels of its window."""
    kind, xw, yw, alpha = film
    x0, y0, w, h = tile
    table = np.array([[_filter64(kind, xw, yw, alpha, (x + 0.5) * xw / 16, (y + 0.5) * yw / 16)
                       for x in range(16)] for y in range(16)])
    acc = np.zeros((h, w, 3))
    wsum = np.zeros((h, w))
    for ly in range(h):
        for lx in range(w):
            for s in range(spp):
                jx, jy = _jitter(orc, seed, width, x0 + lx, y0 + ly, s)
                dx, dy = x0 + lx + jx - 0.5, y0 + ly + jy - 0.5
                for py in range(max(int(np.ceil(dy - yw)), y0), min(int(np.floor(dy + yw)), y0 + h - 1) + 1):
                    for px in range(max(int(np.ceil(dx - xw)), x0), min(int(np.floor(dx + xw)), x0 + w - 1) + 1):
                        ix = min(int(np.floor(abs((px - dx) / xw * 16))), 15)
                        iy = min(int(np.floor(abs((py - dy) / yw * 16))), 15)
                        wt = table[iy, ix]
                        acc[py - y0, px - x0] += wt * samples[(ly * w + lx) * spp + s]
                        wsum[py - y0, px - x0] += wt
    out = np.zeros((h, w, 3))
    nz = wsum > 0
    out[nz] = np.maximum(0.0, acc[nz] / wsum[nz][:, None])
    return out


@pytest.mark.parametrize("film", [(0, 1.5, 1.0, 2.0), (1, 2.0, 1.5, 2.0), (2, 2.0, 2.0, 1.5), (2, 1.0, 0.5, 4.0)])
def test_film_gather_equals_the_reference_splat(orc, film):
    rng = np.random.default_rng(11)
    width, tile, spp, seed = 64, (5, 9, 13, 11), 3, 77
    x0, y0, w, h = tile
    samples = rng.uniform(0.0, 4.0, (w * h * spp, 3)).astype(np.float32)
    got = orc.film(samples, width, tile, spp, seed, film)
    ref = _splat64(orc, samples.astype(np.float64), width, tile, spp, seed, film)
    assert np.allclose(got, ref, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("film", [(0, 2.5, 0.75, 2.0), (1, 1.0, 3.0, 2.0), (2, 2.0, 2.0, 0.5)])
def test_film_of_constant_radiance_is_that_constant(orc, film):
    width, tile, spp, seed = 40, (0, 0, 40, 30), 2, 5
    samples = np.tile(np.array([[0.25, 1.5, 3.0]], np.float32), (40 * 30 * spp, 1))
    got = orc.film(samples, width, tile, spp, seed, film)
    assert np.allclose(got, [0.25, 1.5, 3.0], rtol=1e-6)


def test_film_default_box_is_the_pixel_mean(orc):
    rng = np.random.default_rng(3)
    samples = rng.uniform(0, 1, (6 * 4 * 5, 3)).astype(np.float32)
    got = orc.film(samples, 6, (0, 0, 6, 4), 5, 1)
    s = samples.reshape(4, 6, 5, 3)
    exp = s[:, :, 0]
    for k in range(1, 5):
        exp = exp + s[:, :, k]  # float32, sample order
    assert np.array_equal(got, exp / np.float32(5))


def test_film_gaussian_weights_are_radial(orc):
    # one bright sample in a dark tile: the filtered image is the filter's
    # footprint, symmetric about the sample position
    width, tile, spp, seed = 32, (0, 0, 32, 32), 1, 9
    samples = np.zeros((32 * 32, 3), np.float32)
    c = 16 * 32 + 16
    samples[c] = 1.0
    got = orc.film(samples, width, tile, spp, seed, (2, 2.0, 2.0, 1.0))
    jx, jy = _jitter(orc, seed, width, 16, 16, 0)
    dx, dy = 16 + jx - 0.5, 16 + jy - 0.5
    ys, xs = np.nonzero(got[..., 0])
    assert len(xs) > 4
    assert np.all(np.abs(xs - dx) <= 2.0) and np.all(np.abs(ys - dy) <= 2.0)


# --------------------------------------------------------------- BSDFs ----

def _dir(theta, phi):
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], np.float32)


def test_pow_series_against_float64(orc):
    rng = np.random.default_rng(1)
    xs = rng.uniform(1e-6, 1.0, 3000)
    ys = rng.uniform(0.0, 60.0, 3000)
    for x, y in zip(xs.astype(np.float32), ys.astype(np.float32)):
        exp = float(np.float64(x) ** np.float64(y))
        if exp < 1e-30:
            continue
        # one float rounding of y log2 x dominates: error ~ 2^-24 ln2 |y log2 x|
        assert abs(orc.powf(x, y) / exp - 1.0) < 2.5e-7 * (1.0 + abs(float(y) * np.log2(float(x)))), (x, y)
    assert orc.powf(0.0, 2.0) == 0.0 and orc.powf(0.0, 0.0) == 1.0 and orc.powf(1.0, 123.0) == 1.0
    assert orc.powf(0.5, 3.0) == 0.125 and orc.powf(4.0, 0.5) == pytest.approx(2.0, rel=1e-7)


def _on(sigma_deg, rgb=(0.5, 0.6, 0.7)):
    s = np.float32(min(max(sigma_deg, 0.0), 90.0)) * np.float32(np.pi / 180)
    s2 = s * s
    return (4, rgb, (sigma_deg, 1 - s2 / (2 * (s2 + 0.33)), 0.45 * s2 / (s2 + 0.09)))


def test_orennayar_sigma_zero_is_lambert(orc):
    rng = np.random.default_rng(2)
    ng = np.array([0, 0, 1], np.float32)
    for _ in range(200):
        wo, wi = _dir(rng.uniform(0, 1.5), rng.uniform(0, 6.28)), _dir(rng.uniform(0, 1.5), rng.uniform(0, 6.28))
        a, pa = orc.bsdf_eval(_on(0.0), wo, wi, ng)
        b, pb = orc.bsdf_eval((0, (0.5, 0.6, 0.7)), wo, wi, ng)
        assert np.array_equal(a, b) and pa == pb


def test_orennayar_closed_forms_and_reciprocity(orc):
    m = _on(30.0)
    A = m[2][1]
    ng = np.array([0, 0, 1], np.float32)
    # normal incidence: sin(theta_i) = 0 -> f = rho / pi * A
    f, _ = orc.bsdf_eval(m, _dir(0.7, 1.0), np.array([0, 0, 1], np.float32), ng)
    assert np.allclose(f, np.array(m[1]) / np.pi * A, rtol=1e-6)
    rng = np.random.default_rng(3)
    for _ in range(100):
        a, b = _dir(rng.uniform(0.05, 1.5), rng.uniform(0, 6.28)), _dir(rng.uniform(0.05, 1.5), rng.uniform(0, 6.28))
        assert np.allclose(orc.bsdf_eval(m, a, b, ng)[0], orc.bsdf_eval(m, b, a, ng)[0], rtol=2e-5)
    # rougher surfaces scatter more to grazing back-scatter: f(back) > Lambert's at grazing
    g, _ = orc.bsdf_eval(m, _dir(1.3, 0.0), _dir(1.3, 0.0), ng)
    assert g[0] > m[1][0] / np.pi * A


def _albedo_quadrature(orc, m, wo, n=96):
    ng = np.array([0, 0, 1], np.float32)
    tot = 0.0
    th = (np.arange(n) + 0.5) / n * (np.pi / 2)
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    for t in th:
        for p in ph:
            f, _ = orc.bsdf_eval(m, wo, _dir(t, p), ng)
            tot += f[0] * np.cos(t) * np.sin(t)
    return tot * (np.pi / 2 / n) * (2 * np.pi / (2 * n))


@pytest.mark.parametrize("exponent", [5.0, 60.0])
def test_microfacet_sampling_matches_its_pdf_and_conserves_energy(orc, exponent):
    m = (5, (1.0, 1.0, 1.0), (exponent, 0.0))  # no Fresnel: maximal energy
    ng = np.array([0, 0, 1], np.float32)
    wo = _dir(0.6, 0.3)
    albedo = _albedo_quadrature(orc, m, wo)
    assert 0.3 < albedo <= 1.0 + 1e-3
    rng = np.random.default_rng(4)
    est = []
    for u1, u2 in rng.uniform(0, 1, (20000, 2)).astype(np.float32):
        wi, f, pdf = orc.bsdf_sample(m, wo, ng, u1, u2)
        if pdf > 0 and f[0] > 0:
            _, pdf2 = orc.bsdf_eval(m, wo, wi, ng)
            assert pdf2 == pytest.approx(pdf, rel=2e-3)  # sampling and evaluation agree on the pdf
            est.append(f[0] * abs(wi[2]) / pdf)
        else:
            est.append(0.0)
    est = np.array(est)
    se = est.std() / np.sqrt(len(est))
    assert abs(est.mean() - albedo) < 4 * se + 2e-3
    # reciprocity
    for _ in range(50):
        a, b = _dir(rng.uniform(0.05, 1.4), rng.uniform(0, 6.28)), _dir(rng.uniform(0.05, 1.4), rng.uniform(0, 6.28))
        assert np.allclose(orc.bsdf_eval(m, a, b, ng)[0], orc.bsdf_eval(m, b, a, ng)[0], rtol=1e-4)


def test_glass_fresnel_split_at_normal_incidence(orc):
    m = (3, (1.0, 1.0, 1.0), (1.0, 1.5))
    ng = np.array([0, 0, 1], np.float32)
    R = ((1.5 - 1.0) / 2.5) ** 2
    wo = np.array([0, 0, 1], np.float32)
    us = (np.arange(2000) + 0.5) / 2000
    refl = []
    for u in us.astype(np.float32):
        wi, w, pdf = orc.bsdf_sample(m, wo, ng, u, 0.5)
        refl.append(wi[2] > 0)
        if wi[2] < 0:  # entering: radiance scaled by (1 / 1.5)^2
            assert np.allclose(w, 1 / 2.25, rtol=1e-6) and np.allclose(wi, [0, 0, -1], atol=1e-7)
    assert abs(np.mean(refl) - R) < 1e-3
    # from inside beyond the critical angle: always reflected
    wo_in = _dir(np.pi - 0.9, 0.2)  # 0.9 rad from the inward normal > asin(1/1.5) = 0.73
    for u in (0.01, 0.5, 0.99):
        wi, _, _ = orc.bsdf_sample(m, wo_in, ng, u, 0.5)
        assert wi[2] < 0  # stays inside


def test_glass_pane_transmits_le_t2_over_1_minus_r2(orc):
    # camera -> glass pane (two faces, n = 1.5) -> emitter: radiance
    # Le T^2 / (1 - R^2), T = 1 - R, R = 0.04 at normal incidence (the
    # (1/n)^2 and n^2 radiance scalings cancel through both faces)
    def q(z, flip):
        v = np.array([[-5, -5, z], [5, -5, z], [5, 5, z], [-5, 5, z]], np.float32)
        return v, np.array([[0, 2, 1], [0, 3, 2]] if flip else [[0, 1, 2], [0, 2, 3]], np.uint32)
    front, back = q(-1.0, True), q(-0.9, False)     # outward normals -z and +z
    light = q(0.0, True)                             # normal -z: faces the camera
    v, idx = _merge(front, back, light)
    tri_mat = np.array([0, 0, 0, 0, 1, 1], np.uint32)
    mats = [(3, (1.0, 1.0, 1.0), (1.0, 1.5)), (2, (5.0, 5.0, 5.0))]
    c2 = orc.Bvh(v, idx).compact2()
    cam = ((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 10.0, 24, 24)
    img = orc.render_pt(c2, v, idx, tri_mat, mats, np.array([4, 5], np.uint32), cam, (0, 0, 24, 24), 64, 7, 10, 8,
                        0.5)
    R = 0.04
    exp = 5.0 * (1 - R) ** 2 / (1 - R * R)
    assert abs(img.mean() / exp - 1.0) < 0.01, (img.mean(), exp)


def test_orennayar_sigma_zero_renders_as_lambert(orc):
    v, idx, tm, mats, lights = _box_room()
    c2 = orc.Bvh(v, idx).compact2()
    cam = ((0.0, 1.0, -0.95), (0.0, 1.0, 1.0), (0.0, 1.0, 0.0), 70.0, 16, 16)
    on = [_on(0.0, mats[0][1])] + mats[1:]
    a = orc.render_pt(c2, v, idx, tm, mats, lights, cam, (0, 0, 16, 16), 4, 3, 8, 3, 0.5)
    b = orc.render_pt(c2, v, idx, tm, on, lights, cam, (0, 0, 16, 16), 4, 3, 8, 3, 0.5)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_bdpt_agrees_with_pt_on_glossy_and_rough_materials(orc):
    # the box with an Oren-Nayar floor and a conductor microfacet left wall:
    # connections now evaluate glossy BSDFs at both ends
    v, idx, tm, mats, lights = _box_room()
    tm = tm.copy()
    tm[0:2] = 3   # floor
    tm[8:10] = 4  # left wall
    mats = mats + [_on(30.0, (0.7, 0.6, 0.5)), (5, (0.9, 0.8, 0.6), (25.0, 1.0, 0.2, 0.4, 1.4, 3.6, 2.4, 1.8))]
    c2 = orc.Bvh(v, idx).compact2()
    W = H = 20
    cam = ((0.0, 1.0, -0.95), (0.0, 1.0, 1.0), (0.0, 1.0, 0.0), 70.0, W, H)
    pt, bd = [], []
    for seed in range(8):
        pt.append(orc.render_pt(c2, v, idx, tm, mats, lights, cam, (0, 0, W, H), 512, seed, 24, 3, 0.5)
                  .reshape(-1, 3).mean(0))
        bd.append(orc.render_bdpt(c2, v, idx, tm, mats, lights, cam, (0, 0, W, H), 512, seed, 24, 3, 0.5, 4096, 4, 3)
                  .reshape(-1, 3).mean(0))
    pt, bd = np.mean(pt, 0), np.mean(bd, 0)
    assert np.all(pt > 0.05)
    assert np.all(np.abs(bd / pt - 1.0) < 0.015), (pt, bd)


# ------------------------------------------------------- ray-cast debug ----

# kernel_ray_cast_update's colour table (src/integrators/integrator_raycast.cu:47-61)
RAYCAST_PALETTE = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 1, 1], [1, 0, 1], [1, 1, 0], [0.1, 0.1, 0.1],
                            [0.5, 0.5, 0.5], [1, 1, 1], [0, 0.25, 0.5], [1, 0.5, 0], [0.5, 0.25, 0],
                            [0.2, 0.9, 1.0], [0.7, 0.3, 0.1], [0.34, 0.89, 0.45]], np.float32)


def _open_room():
    """The box room without its front wall and ceiling, seen from outside:
    some camera rays miss everything."""
    v, idx, tm, mats, _ = _box_room()
    keep = np.r_[0:2, 4:6, 8:12, 12:14]  # floor, back, left, right, light
    mats = mats + [(5, (0.9, 0.8, 0.6), (25.0, 1.0, 0.2, 0.4, 1.4, 3.6, 2.4, 1.8)), (3, (0.8, 1.0, 0.9), (1.0, 1.5))]
    tm = tm[keep].copy()
    tm[2:4] = 3  # back wall: microfacet
    tm[4:6] = 4  # left wall: glass
    return v, idx[keep], tm, mats


RAYCAST_CAM = ((0.3, 2.6, -3.0), (0.0, 0.8, 0.5), (0.0, 1.0, 0.0), 60.0, 24, 20)


@pytest.mark.parametrize("shade", range(5))
def test_raycast_colours_the_traced_hits(orc, shade):
    # one sample per pixel, box film: each pixel is the colour of its camera
    # ray's closest hit, traced independently here
    v, idx, tm, mats = _open_room()
    c2 = orc.Bvh(v, idx).compact2()
    W, H = RAYCAST_CAM[4], RAYCAST_CAM[5]
    img = orc.render_raycast(c2, v, idx, tm, mats, RAYCAST_CAM, (0, 0, W, H), 1, 5, shade).reshape(-1, 3)
    rays, ids, t = _primary(orc, c2, RAYCAST_CAM, 1, 5)
    hit = ids >= 0
    assert 0 < hit.sum() < len(ids)  # both hits and misses
    assert np.all(img[~hit] == 0.0)
    h = np.nonzero(hit)[0]
    tri = idx[ids[h]].astype(np.int64)
    v0, v1, v2 = (v[tri[:, k]].astype(np.float64) for k in range(3))
    if shade == 0:
        assert np.array_equal(img[h], RAYCAST_PALETTE[ids[h] % 15])
    elif shade == 1:
        cols = np.array([m[1] for m in mats], np.float32)
        assert np.array_equal(img[h], cols[tm[ids[h]]])
    elif shade == 2:
        n = np.cross(v1 - v0, v2 - v0)
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        assert np.allclose(img[h], n, atol=1e-6)
    elif shade == 3:
        u, w_ = img[h, 0].astype(np.float64), img[h, 1].astype(np.float64)
        assert np.all(img[h, 2] == 0.0)
        assert np.all((u > -1e-5) & (w_ > -1e-5) & (u + w_ < 1 + 1e-5))
        p = rays[h, :3].astype(np.float64) + rays[h, 4:7].astype(np.float64) * t[h, None].astype(np.float64)
        q = u[:, None] * v0 + w_[:, None] * v1 + (1 - u - w_)[:, None] * v2
        assert np.allclose(q, p, atol=2e-5)
    else:
        assert np.all(img[h] == 1.0)


def test_raycast_uniform_is_the_hit_fraction(orc):
    # uniform shading with spp samples per pixel: the pixel mean is the
    # fraction of its camera rays that hit anything
    v, idx, tm, mats = _open_room()
    c2 = orc.Bvh(v, idx).compact2()
    W, H = RAYCAST_CAM[4], RAYCAST_CAM[5]
    spp = 6
    img = orc.render_raycast(c2, v, idx, tm, mats, RAYCAST_CAM, (0, 0, W, H), spp, 9, 4)
    _, ids, _ = _primary(orc, c2, RAYCAST_CAM, spp, 9)
    frac = (ids >= 0).reshape(H, W, spp).mean(axis=2)
    assert np.allclose(img[..., 0], frac, atol=1e-6)
    assert np.any((frac > 0) & (frac < 1))  # silhouette pixels are partly covered


# ------------------------------------------------------------- furnace ----

def _icosphere(sub):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(sub):
        cache, nf = {}, []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[k] = len(v) - 1
            return cache[k]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v, np.float32), np.array(f, np.uint32)


def _furnace(albedo, le):
    """Inside a sphere every surface element sees every other with a form
    factor proportional to its area, so with a non-reflecting emitter cap of
    area fraction f (radiance Le) and Lambert walls of albedo a the walls'
    radiance solves L = a (f Le + (1 - f) L): L = a f Le / (1 - a (1 - f))."""
    v, idx = _icosphere(3)
    c = v[idx].mean(1)
    n = np.cross(v[idx[:, 1]] - v[idx[:, 0]], v[idx[:, 2]] - v[idx[:, 0]])
    flip = (n * c).sum(1) > 0  # wind every face to face the centre
    idx[flip] = idx[flip][:, [0, 2, 1]]
    cap = c[:, 1] > 0.75
    area = 0.5 * np.linalg.norm(np.cross(v[idx[:, 1]] - v[idx[:, 0]], v[idx[:, 2]] - v[idx[:, 0]]), axis=1)
    f = area[cap].sum() / area.sum()
    tm = cap.astype(np.uint32)
    mats = [(0, (albedo,) * 3), (2, (le,) * 3)]
    lights = np.nonzero(cap)[0].astype(np.uint32)
    cam = ((0.0, 0.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), 60.0, 16, 16)  # looks at the walls only
    return v, idx, tm, mats, lights, cam, albedo * f * le / (1 - albedo * (1 - f))


@pytest.mark.parametrize("albedo", [0.5, 0.8])
def test_pt_spherical_furnace_closed_form(orc, albedo):
    # every bounce of the multi-bounce estimator (NEE + MIS, roulette, depth)
    # enters the geometric series; the faceted sphere is within 0.1% of round
    v, idx, tm, mats, lights, cam, exp = _furnace(albedo, 4.0)
    c2 = orc.Bvh(v, idx).compact2()
    got = np.mean([orc.render_pt(c2, v, idx, tm, mats, lights, cam, (0, 0, 16, 16), 64, s, 60, 3, 0.5).mean()
                   for s in range(4)])
    assert abs(got / exp - 1) < 0.004, (got, exp)


def test_bdpt_spherical_furnace_closed_form(orc):
    v, idx, tm, mats, lights, cam, exp = _furnace(0.5, 4.0)
    c2 = orc.Bvh(v, idx).compact2()
    got = np.mean([orc.render_bdpt(c2, v, idx, tm, mats, lights, cam, (0, 0, 16, 16), 64, s, 24, 3, 0.5, 4096, 6,
                                   3).mean() for s in range(4)])
    assert abs(got / exp - 1) < 0.006, (got, exp)


def test_bdpt_agrees_with_pt_under_shading_normals(orc):
    # the box with vertex normals tilted ~20 degrees off the floor's and the
    # back wall's geometric normals: camera paths shade with them, light
    # paths too, with the adjoint factor |wo.ns||wi.ng| / (|wo.ng||wi.ns|)
    # importance transport needs there (at the light pass's bounces and at
    # the cache vertex of each connection); without that factor BDPT and the
    # path tracer estimate different images
    v, idx, tm, mats, lights = _box_room()
    nt = len(idx)
    tilt = np.array([[0.35, 1.0, 0.3], [-0.3, 1.0, 0.35], [0.3, 1.0, -0.35], [-0.35, 1.0, -0.3]], np.float32)
    tilt /= np.linalg.norm(tilt, axis=1, keepdims=True)
    back = tilt[:, [0, 2, 1]].copy()   # about +z for the back wall
    normals = np.concatenate([tilt, back])
    nidx = np.full((nt, 3), 0xFFFFFFFF, np.uint32)
    nidx[0:2] = [[0, 1, 2], [0, 2, 3]]   # floor corners 0..3
    nidx[4:6] = [[4, 5, 6], [4, 6, 7]]   # back wall
    nrm = orc.tri_normals(normals, nidx, nt)
    c2 = orc.Bvh(v, idx).compact2()
    W = H = 20
    cam = ((0.0, 1.0, -0.95), (0.0, 1.0, 1.0), (0.0, 1.0, 0.0), 70.0, W, H)
    pt, bd = [], []
    for seed in range(8):
        pt.append(orc.render_pt(c2, v, idx, tm, mats, lights, cam, (0, 0, W, H), 512, seed, 24, 3, 0.5,
                                normals=nrm).reshape(-1, 3).mean(0))
        bd.append(orc.render_bdpt(c2, v, idx, tm, mats, lights, cam, (0, 0, W, H), 512, seed, 24, 3, 0.5, 4096, 4, 3,
                                  normals=nrm).reshape(-1, 3).mean(0))
    pt, bd = np.mean(pt, 0), np.mean(bd, 0)
    flat = orc.render_pt(c2, v, idx, tm, mats, lights, cam, (0, 0, W, H), 512, 0, 24, 3, 0.5).reshape(-1, 3).mean(0)
    assert np.all(np.abs(pt / flat - 1.0) > 0.01), (pt, flat)  # the normals change the image
    assert np.all(np.abs(bd / pt - 1.0) < 0.015), (pt, bd)
