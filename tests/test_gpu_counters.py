"""GPU tests of the render passes' measurement counters against the oracle:
rb_render_params.path_counter[0] is the number of camera rays whose closest
hit is a non-emitter (the vertices shaded at bounce 0), counted from the
oracle's camera rays and closest hits; later bounces never hold more paths
than the one before; BDPT's light_path_counter[0] is the number of light
sub-paths (every one is emitted when the scene has emitters); ray_counter
is the sum of what the passes traced and at least the camera rays."""
import ctypes as C

import numpy as np
import pytest

from renderbox2_amd import _lib as L

pytestmark = pytest.mark.gpu


def _c2(bvh):
    n, w, x = bvh.export(L.RB_LAYOUT_REF_COMPACT2)
    return n.reshape(-1, 4), w.reshape(-1, 4), x


@pytest.mark.parametrize("kind", ["pt", "bdpt"])
def test_path_counters_match_the_oracle_camera_hits(gpu, orc, kind):
    s = gpu.scenes.pt_room(spheres_hi=10, spheres_lo=50)
    b = gpu.scenes.build(s)
    tr = gpu.CudaTracer(b, s)
    W, H, spp, seed, depth = 48, 40, 2, 7, 6
    cam = gpu.scenes.pt_camera(W, H)
    params = dict(spp=spp, seed=seed, max_depth=depth, rr_start=3, rr_prob=0.5)
    if kind == "bdpt":
        params.update(light_paths=3000, light_max_depth=4, connections=3)
    r = gpu.SamplerRenderer(tr, cam, kind, **params)
    paths = gpu.DeviceBuffer(8 * (depth + 1))
    paths.zero()
    lpaths = gpu.DeviceBuffer(8 * 4)
    lpaths.zero()
    rays = gpu.DeviceBuffer(8)
    rays.zero()
    r.params.path_counter = C.c_void_p(paths.ptr)
    r.params.light_path_counter = C.c_void_p(lpaths.ptr)
    r.params.ray_counter = C.c_void_p(rays.ptr)
    r.render()
    got = paths.download(np.uint64, depth + 1).astype(np.int64)
    # oracle: camera rays of the frame and their closest hits
    c = cam.c
    r2c, c2w = orc.camera_matrices(tuple(c.position), tuple(c.lookat), tuple(c.up), c.hfov_deg, W, H)
    prim = orc.raygen_camera(r2c, c2w, W, (0, 0, W, H), spp, seed)
    ids, _ = orc.trace_compact2(*_c2(b), prim)
    _, _, tm, _ = s.geometry()
    mats = s.materials()
    hit = ids >= 0
    emitter = np.zeros(len(ids), bool)
    emitter[hit] = [mats[m][0] == L.RB_MAT_EMITTER for m in tm[ids[hit]]]
    assert got[0] == int((hit & ~emitter).sum()) > 0
    assert np.all(np.diff(got) <= 0)  # paths only end
    assert got[-1] >= 0
    traced = int(rays.download(np.uint64, 1)[0])
    assert traced >= W * H * spp + got[0]  # camera rays, then at least a light sample or bounce per shaded vertex
    if kind == "bdpt":
        lgot = lpaths.download(np.uint64, 4).astype(np.int64)
        assert lgot[0] == 3000
        assert np.all(np.diff(lgot) <= 0) and lgot[1] > 0
