"""GPU tests of the request server: results equal the same request rendered
alone, bit for bit, whatever else shares the GPU and in both serving modes
(each request on its own stream, or all in lockstep with one batched trace
per round); admission within the memory budget; rejection, injected
allocation failure and busy scenes."""
import numpy as np
import pytest

from renderbox2_amd import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(gpu):
    room = gpu.scenes.pt_room(spheres_hi=10, spheres_lo=50)
    tr_room = gpu.CudaTracer(gpu.scenes.build(room), room)
    ao = gpu.scenes.ao_room()
    tr_ao = gpu.CudaTracer(gpu.scenes.build(ao, max_leaf=4), ao)
    return tr_room, tr_ao


def _mix(gpu, ids):
    room, ao = ids
    pc, ac = gpu.scenes.pt_camera(64, 48), gpu.scenes.ao_camera(64, 64)
    return [
        ("pt", room, pc, None, dict(spp=2, seed=1, max_depth=6)),
        ("ao", ao, ac, None, dict(spp=2, seed=2, ao_samples=8)),
        ("bdpt", room, pc, None, dict(spp=1, seed=3, max_depth=6, light_paths=2048, light_max_depth=3,
                                      connections=2)),
        ("pt", room, pc, (16, 8, 32, 24), dict(spp=4, seed=4, max_depth=8)),
        ("ao", room, pc, (0, 0, 32, 48), dict(spp=1, seed=5, ao_samples=4, ao_radius=2.0)),
        ("bdpt", room, pc, (32, 0, 32, 48), dict(spp=2, seed=6, light_paths=1024, light_max_depth=4,
                                               connections=4)),
        ("pt", ao, ac, None, dict(spp=1, seed=7)),
        ("ao", ao, ac, (8, 8, 40, 40), dict(spp=3, seed=8, ao_samples=2)),
        ("raycast", room, pc, (8, 8, 40, 32), dict(spp=2, seed=9, shade=L.RB_SHADE_MATERIAL_ID)),
    ]


def _alone(gpu, tracers, kind, cam, tile, params):
    return gpu.SamplerRenderer(tracers, cam, kind, **params).render(tile)


@pytest.mark.parametrize("mode", ["streams", "lockstep"])
def test_served_results_equal_standalone_renders(gpu, scenes, mode):
    srv = gpu.RenderServer(max_inflight=4, mode=mode)
    ids = (srv.add_scene(scenes[0]), srv.add_scene(scenes[1]))
    tr = {ids[0]: scenes[0], ids[1]: scenes[1]}
    mix = _mix(gpu, ids) * 2  # every request twice: the second copy runs beside different neighbours
    rids = [srv.submit_render(k, sc, cam, tile, **p) for k, sc, cam, tile, p in mix]
    for rid, (k, sc, cam, tile, p) in zip(rids, mix):
        st, img = srv.wait(rid)
        assert st.state == L.RB_REQ_DONE and st.result == L.RB_OK
        assert st.rays > 0 and st.latency_ms >= st.queue_ms >= 0.0
        ref = _alone(gpu, tr[sc], k, cam, tile, p)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (k, tile, p)
    s = srv.stats()
    assert s["completed"] == len(mix) and s["failed"] == 0
    assert s["p99_ms"] >= s["p50_ms"] > 0.0
    # every scratch allocation came from the slots' arenas (the estimate holds)
    assert s["arena_misses"] == 0 and s["peak_arena_bytes"] > 0
    srv.close()


@pytest.mark.parametrize("mode", ["streams", "lockstep"])
def test_trace_requests_match_rb_trace(gpu, scenes, mode):
    tr = scenes[0]
    n = 50000
    rng = np.random.default_rng(3)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(-4.0, 4.0, (n, 3))  # inside the room (walls at +-5)
    rays[:, 4:7] = d
    rays[:, 7] = 1e30
    ids, t = tr.trace(rays)
    srv = gpu.RenderServer(max_inflight=2, mode=mode)
    sid = srv.add_scene(tr)
    dr = gpu.DeviceBuffer.from_array(rays)
    dh = [gpu.DeviceBuffer(n * 8) for _ in range(3)]
    rids = [srv.submit_trace(sid, dr, h, n, L.RB_ANY if i == 2 else L.RB_CLOSEST) for i, h in enumerate(dh)]
    for i, (rid, h) in enumerate(zip(rids, dh)):
        st, _ = srv.wait(rid)
        assert st.result == L.RB_OK and st.rays == n
        got = h.download(np.int32, 2 * n).reshape(n, 2)
        if i < 2:
            assert np.array_equal(got[:, 0], ids) and np.array_equal(got[:, 1], t.view(np.int32))
        else:
            assert np.array_equal(got[:, 0] >= 0, ids >= 0)  # any-hit: same occlusion answers
    srv.close()


@pytest.mark.parametrize("mode", ["streams", "lockstep"])
def test_small_budget_defers_but_completes(gpu, scenes, mode):
    big = gpu.RenderServer.render_request("pt", 0, gpu.scenes.pt_camera(64, 48), None, spp=2, seed=1, max_depth=6)
    need = gpu.RenderServer.request_bytes(big)
    srv = gpu.RenderServer(max_inflight=8, hbm_budget=int(need * 1.5), mode=mode)  # one at a time
    sid = srv.add_scene(scenes[0])
    cam = gpu.scenes.pt_camera(64, 48)
    rids = [srv.submit_render("pt", sid, cam, None, spp=2, seed=s, max_depth=6) for s in range(6)]
    imgs = [srv.wait(r)[1] for r in rids]
    s = srv.stats()
    assert s["completed"] == 6 and s["deferrals"] > 0
    assert s["peak_budget_bytes"] <= int(need * 1.5)
    for sd, img in enumerate(imgs):
        ref = _alone(gpu, scenes[0], "pt", cam, None, dict(spp=2, seed=sd, max_depth=6))
        assert np.array_equal(img, ref)
    srv.close()


def test_request_larger_than_budget_is_rejected(gpu, scenes):
    srv = gpu.RenderServer(hbm_budget=1 << 20)
    sid = srv.add_scene(scenes[0])
    with pytest.raises(gpu.RbError) as e:
        srv.submit_render("pt", sid, gpu.scenes.pt_camera(256, 256), None, spp=4)
    assert e.value.status == L.RB_ERR_OOM
    assert srv.stats()["rejected"] == 1
    with pytest.raises(gpu.RbError):
        srv.submit_render("pt", sid + 7, gpu.scenes.pt_camera(8, 8))  # unknown scene
    srv.close()


@pytest.mark.parametrize("mode", ["streams", "lockstep"])
def test_injected_allocation_failure_fails_one_request_only(gpu, scenes, mode):
    srv = gpu.RenderServer(max_inflight=1, mode=mode)
    sid = srv.add_scene(scenes[0])
    cam = gpu.scenes.pt_camera(32, 32)
    srv.inject_alloc_failures(1)
    r1 = srv.submit_render("ao", sid, cam, None, spp=1, seed=1)
    r2 = srv.submit_render("ao", sid, cam, None, spp=1, seed=2)
    st1, _ = srv.wait(r1)
    st2, img2 = srv.wait(r2)
    assert st1.state == L.RB_REQ_FAILED and st1.result == L.RB_ERR_OOM
    assert st2.state == L.RB_REQ_DONE
    assert np.array_equal(img2, _alone(gpu, scenes[0], "ao", cam, None, dict(spp=1, seed=2)))
    s = srv.stats()
    assert s["failed"] == 1 and s["completed"] == 1
    srv.close()


def test_scene_in_use_cannot_be_removed(gpu, scenes):
    srv = gpu.RenderServer(max_inflight=1)
    sid = srv.add_scene(scenes[0])
    cam = gpu.scenes.pt_camera(64, 64)
    rids = [srv.submit_render("pt", sid, cam, None, spp=2, seed=s) for s in range(3)]
    with pytest.raises(gpu.RbError) as e:
        srv.remove_scene(sid)
    assert e.value.status == L.RB_ERR_BUSY
    srv.drain()
    for r in rids:
        assert srv.poll(r).state == L.RB_REQ_DONE
    srv.remove_scene(sid)
    srv.close()


@pytest.mark.parametrize("policy", ["deadline", "fifo"])
def test_deadline_policy_lets_short_requests_overtake(gpu, scenes, policy):
    import time
    srv = gpu.RenderServer(max_inflight=1, policy=policy)
    sid = srv.add_scene(scenes[0])
    cam = gpu.scenes.pt_camera(256, 256)
    t_sub, rids = [], []
    for s in range(3):
        t_sub.append(time.perf_counter())
        rids.append(srv.submit_render("pt", sid, cam, None, spp=8, seed=s))
    t_sub.append(time.perf_counter())
    rids.append(srv.submit_render("ao", sid, cam, (0, 0, 32, 32), spp=1, seed=9))
    done = []
    for t0, r in zip(t_sub, rids):
        st, _ = srv.wait(r)
        assert st.result == L.RB_OK
        done.append(t0 + st.latency_ms * 1e-3)
    if policy == "deadline":
        assert done[3] < done[1] and done[3] < done[2]  # admitted ahead of the queued long ones
    else:
        assert done[3] > done[2]  # arrival order
    srv.close()
