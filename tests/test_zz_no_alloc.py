"""rb_trace_fill_noalloc (the allocation-free variant of rb_trace_fill)
against rb_trace_fill, through the C-ABI: two calls in a row, the second
with fewer rays, every field of the intersection records bit for bit; a
call over its size limit is refused; the misses' fields keep what the
caller had there, as with rb_trace_fill."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_rays

pytestmark = pytest.mark.gpu

FIELDS = [("position", 3, np.float32), ("geometric_normal", 3, np.float32), ("shading_normal", 3, np.float32),
          ("incoming_direction", 3, np.float32), ("uv", 1, np.float32), ("primitive_id", 1, np.uint32),
          ("material_id", 1, np.uint32), ("front", 1, np.uint8), ("intersected", 1, np.uint8),
          ("epsilon", 1, np.float32), ("t", 1, np.float32)]
NOALLOC_MAX = 1 << 20


def _soa(gpu, n):
    """An rb_isect_soa over fresh device buffers (filled with 0xAB so the
    entries a call leaves alone compare equal too)."""
    from renderbox2_amd import _lib as L
    soa = L.rb_isect_soa()
    bufs = {}
    for name, k, dt in FIELDS:
        count = 2 * n if name == "uv" else n
        for j in range(k):
            b = gpu.DeviceBuffer(max(count, 1) * np.dtype(dt).itemsize)
            b.upload(np.full(b.nbytes, 0xAB, np.uint8))
            bufs[(name, j)] = (b, count, dt)
            if k == 3:
                getattr(soa, name)[j] = b.ptr
            else:
                setattr(soa, name, b.ptr)
    return soa, bufs


def _read(bufs):
    return {key: b.download(dt, count) for key, (b, count, dt) in bufs.items()}


def test_noalloc_fill_equals_trace_fill(gpu):
    from renderbox2_amd import _lib as L
    lib = L.lib()
    s = gpu.scenes.ao_room(seed=1)
    tr = gpu.CudaTracer(gpu.scenes.build(s, max_leaf=4), s)
    for n, seed in ((NOALLOC_MAX, 11), (70_001, 12)):  # full size, then a smaller call reusing the scratch
        rays = random_rays(n, seed=seed, box=4.5, tmax=2.0)  # inside the closed room: short rays miss
        d_rays = gpu.DeviceBuffer.from_array(rays)
        for flags in (L.RB_CLOSEST, L.RB_ANY):
            ref_soa, ref_bufs = _soa(gpu, n)
            d_hits = gpu.DeviceBuffer(n * 8)
            L.check(lib.rb_trace_fill(tr.handle, d_rays.ptr, n, flags, None, d_hits.ptr, C.byref(ref_soa), None))
            soa, bufs = _soa(gpu, n)
            L.check(lib.rb_trace_fill_noalloc(tr.handle, d_rays.ptr, n, flags, None, C.byref(soa), None))
            L.check(lib.rb_stream_synchronize(None))
            ref, got = _read(ref_bufs), _read(bufs)
            hit = ref[("intersected", 0)] == 1
            assert 0.2 < hit.mean() < 0.95
            for key in ref:
                assert np.array_equal(got[key].view(np.uint8), ref[key].view(np.uint8)), (n, flags, key)
    # over the limit: refused before anything runs
    st = lib.rb_trace_fill_noalloc(tr.handle, d_rays.ptr, NOALLOC_MAX + 1, 0, None, C.byref(soa), None)
    assert st == L.RB_ERR_INVALID and "at most" in L.last_error()
