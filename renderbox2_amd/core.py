"""Host-side mirror of renderbox2's scene / BVH / tracer interface over the C-ABI.

Class and method names follow the reference (paths relative to renderbox2/):
  Scene            src/core/scene.h:55-199
  Platform         src/accelerators/sbvh/platform.h:78-135
  BVH, BuildParams src/accelerators/sbvh/bvh.h:75-133
  CudaBvh          src/accelerators/sbvh/cudabvh.h:64-120
  CudaTracer       src/accelerators/sbvh/cudatracer.h:63-117
  PerspectiveCamera src/core/camera.h:39-91
Every computation runs in librb.so (host C++ for the SBVH build, HIP kernels
for everything per ray); this module only moves arguments.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib, ptr

RAY_DTYPE = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"),
                      ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("id", "<i4"), ("t", "<f4")])
FLT_MAX = np.float32(3.4028234663852886e38)


def device_init(device=0):
    check(lib().rb_device_init(device))


class DeviceBuffer:
    """A device allocation (MemoryAllocator::allocate, src/memory/memoryallocator.cpp:54-67)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().rb_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().rb_memcpy_htod(b.ptr, ptr(a), a.nbytes, stream))
        return b

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().rb_memcpy_htod(self.ptr, ptr(a), a.nbytes, stream))

    def download(self, dtype, count=None, stream=None):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype)
        check(lib().rb_memcpy_dtoh(ptr(out), self.ptr, out.nbytes, stream))
        return out

    def zero(self, stream=None):
        check(lib().rb_memset(self.ptr, 0, self.nbytes, stream))

    def free(self):
        if self.ptr:
            check(lib().rb_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().rb_free(self.ptr)
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(lib().rb_stream_create(C.byref(s)))
        self.handle = s.value

    def synchronize(self):
        check(lib().rb_stream_synchronize(self.handle))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().rb_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(lib().rb_event_create(C.byref(e)))
        self.handle = e.value

    def record(self, stream=None):
        check(lib().rb_event_record(self.handle, stream.handle if isinstance(stream, Stream) else stream))

    def elapsed_ms(self, later):
        ms = C.c_float()
        check(lib().rb_event_elapsed_ms(self.handle, later.handle, C.byref(ms)))
        return ms.value

    def synchronize(self):
        check(lib().rb_event_synchronize(self.handle))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().rb_event_destroy(self.handle)
        except Exception:
            pass


class Scene:
    """Host scene: vertices, triangles, materials, lights (Scene, src/core/scene.h)."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().rb_scene_create(C.byref(h)))
        self._h = h.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().rb_scene_destroy(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def add_material(self, kind, color, param=()):
        m = L.rb_material(kind, (C.c_float * 3)(*color), (C.c_float * 8)(*param))
        mid = C.c_uint32()
        check(lib().rb_scene_add_material(self._h, C.byref(m), C.byref(mid)))
        return mid.value

    def add_glass_bsdf(self, color, etai=1.0, etat=1.5):
        return self.add_material(L.RB_MAT_GLASS, color, (etai, etat))

    def add_orennayar_bsdf(self, color, sigma_deg):
        return self.add_material(L.RB_MAT_ORENNAYAR, color, (sigma_deg,))

    def add_microfacet_bsdf(self, color, exponent, fresnel=None):
        """fresnel: None, ("conductor", eta_rgb, k_rgb) or ("dielectric", etai, etat)."""
        if fresnel is None:
            p = (exponent, L.RB_FRESNEL_NONE)
        elif fresnel[0] == "conductor":
            p = (exponent, L.RB_FRESNEL_CONDUCTOR, *fresnel[1], *fresnel[2])
        else:
            p = (exponent, L.RB_FRESNEL_DIELECTRIC, fresnel[1], fresnel[2])
        return self.add_material(L.RB_MAT_MICROFACET, color, p)

    # Scene::add_diffuse_bsdf / add_mirror_bsdf / add_diffuse_light_bsdf
    def add_diffuse_bsdf(self, albedo):
        return self.add_material(L.RB_MAT_LAMBERT, albedo)

    def add_mirror_bsdf(self, color=(1.0, 1.0, 1.0)):
        return self.add_material(L.RB_MAT_MIRROR, color)

    def add_diffuse_light_bsdf(self, radiance):
        return self.add_material(L.RB_MAT_EMITTER, radiance)

    def add_layered_material(self, layer_ids):
        """A material of 1..4 layers, each a material id of this scene that is
        neither an emitter nor layered (Material's layers,
        src/core/material.h:44-52): the equal mixture of the layers, one
        picked uniformly per sampled direction."""
        ids = np.ascontiguousarray(layer_ids, dtype=np.uint32).ravel()
        mid = C.c_uint32()
        check(lib().rb_scene_add_layered_material(self._h, ids.ctypes.data_as(C.c_void_p), len(ids), C.byref(mid)))
        return mid.value

    def load_obj(self, path, material_id=0):
        """Every object of a Wavefront OBJ file with one material (ObjFileReader, src/io/objfilereader.cpp)."""
        check(lib().rb_scene_load_obj(self._h, str(path).encode(), material_id))

    def load_xml(self, path, width=512, height=512):
        """The reference's XML scene file (CustomFileReader, src/io/customfilereader.cpp):
        mesh, materials, lights; returns the file's camera for a width x height film."""
        c = L.rb_camera()
        check(lib().rb_scene_load_xml(self._h, str(path).encode(), width, height, C.byref(c)))
        cam = PerspectiveCamera.__new__(PerspectiveCamera)
        cam.c = c
        return cam

    def add_mesh(self, verts, idx, material_id, normals=None, nidx=None, texcoords=None, tidx=None):
        """Scene::add_mesh (src/core/scene.cpp:83-123): triangles index the
        mesh's own vertices; normals / texcoords with their own index triples
        (None: one per vertex, indexed like the vertices; a triple of
        0xFFFFFFFF: that triangle has none)."""
        v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
        if normals is None and texcoords is None:
            check(lib().rb_scene_add_mesh(self._h, ptr(v), len(v), ptr(i), len(i), material_id))
            return
        n = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        ni = None if nidx is None else np.ascontiguousarray(nidx, np.uint32).reshape(-1, 3)
        t = None if texcoords is None else np.ascontiguousarray(texcoords, np.float32).reshape(-1, 2)
        ti = None if tidx is None else np.ascontiguousarray(tidx, np.uint32).reshape(-1, 3)
        check(lib().rb_scene_add_mesh_ex(self._h, ptr(v), len(v), ptr(i), len(i),
                                         ptr(n) if n is not None else None, 0 if n is None else len(n),
                                         ptr(ni) if ni is not None else None,
                                         ptr(t) if t is not None else None, 0 if t is None else len(t),
                                         ptr(ti) if ti is not None else None, material_id))

    def shading(self):
        """(normals f32[nn,3], nidx u32[nt,3] or None, texcoords f32[ntc,2],
        tidx u32[nt,3] or None): the scene's shading data."""
        nn, ntc, hn, ht = C.c_uint32(), C.c_uint32(), C.c_int(), C.c_int()
        check(lib().rb_scene_shading_counts(self._h, C.byref(nn), C.byref(ntc), C.byref(hn), C.byref(ht)))
        nt = self.counts()[1]
        n = np.empty((nn.value, 3), np.float32)
        t = np.empty((ntc.value, 2), np.float32)
        ni = np.empty((nt, 3), np.uint32) if hn.value else None
        ti = np.empty((nt, 3), np.uint32) if ht.value else None
        check(lib().rb_scene_shading(self._h, ptr(n), ptr(ni) if ni is not None else None, ptr(t),
                                     ptr(ti) if ti is not None else None))
        return n, ni, t, ti

    def generate(self, kind, count=0, subdiv=0, seed=0, r=(0.0, 0.0), c=(0.0, 0.0), material_id=-1,
                 mix_materials=False, emit=(0.0, 0.0, 0.0), subdiv_hi=0, hi_count=0, smooth=False):
        p = L.rb_gen_params(kind, count, subdiv, seed, r[0], r[1], c[0], c[1], material_id,
                            int(mix_materials), (C.c_float * 3)(*emit), subdiv_hi, hi_count, int(smooth))
        check(lib().rb_scene_generate(self._h, C.byref(p)))

    def counts(self):
        nv, nt, nl, nm = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().rb_scene_counts(self._h, C.byref(nv), C.byref(nt), C.byref(nl), C.byref(nm)))
        return nv.value, nt.value, nl.value, nm.value

    def get_num_triangles(self):
        return self.counts()[1]

    def get_num_lights(self):
        return self.counts()[2]

    def geometry(self):
        """(verts f32[nv,3], idx u32[nt,3], tri_material u32[nt], light_tris u32[nl])"""
        nv, nt, nl, _ = self.counts()
        v = np.empty((nv, 3), np.float32)
        i = np.empty((nt, 3), np.uint32)
        m = np.empty(nt, np.uint32)
        lt = np.empty(nl, np.uint32)
        check(lib().rb_scene_geometry(self._h, ptr(v), ptr(i), ptr(m), ptr(lt)))
        return v, i, m, lt

    def materials(self):
        nm = self.counts()[3]
        arr = (L.rb_material * nm)()
        if nm:
            check(lib().rb_scene_materials(self._h, arr))
        return [(m.kind, tuple(m.color), tuple(m.param)) for m in arr]

    def get_scene_aabb(self):
        lo = np.empty(3, np.float32)
        hi = np.empty(3, np.float32)
        check(lib().rb_scene_bounds(self._h, ptr(lo), ptr(hi)))
        return lo, hi


class Platform:
    """SAH costs and leaf preferences (Platform, src/accelerators/sbvh/platform.h:78-135)."""

    def __init__(self, node_cost=1.0, tri_cost=1.0):
        self.node_cost = node_cost
        self.tri_cost = tri_cost
        self.min_leaf = 1
        self.max_leaf = 0x7FFFFFF

    def set_leaf_preferences(self, min_size, max_size):
        self.min_leaf, self.max_leaf = min_size, max_size


class BuildParams:
    def __init__(self, split_alpha=1.0e-5, threads=0):
        self.split_alpha = split_alpha
        self.threads = threads


class BVH:
    """SBVH built on the host (BVH::BVH, src/accelerators/sbvh/splitbvhbuilder.cpp:177-214)."""

    @staticmethod
    def _params(platform, params):
        platform = platform or Platform()
        params = params or BuildParams()
        return L.rb_build_params(platform.min_leaf, platform.max_leaf, params.split_alpha,
                                 platform.node_cost, platform.tri_cost, params.threads)

    @classmethod
    def _wrap(cls, handle):
        b = cls.__new__(cls)
        b._h = handle
        return b

    @classmethod
    def load_binary(cls, path, scene, platform=None, params=None):
        """The tree saved by save_binary, checked against this scene and these
        build parameters (rb_bvh_load_binary)."""
        bp = cls._params(platform, params)
        h = C.c_void_p()
        check(lib().rb_bvh_load_binary(str(path).encode(), scene.handle, C.byref(bp), C.byref(h)))
        return cls._wrap(h.value)

    @classmethod
    def build_cached(cls, scene, cache_dir, platform=None, params=None):
        """(bvh, from_cache): load <cache_dir>/rbvh_<key>.bin or build and write it."""
        bp = cls._params(platform, params)
        h = C.c_void_p()
        hit = C.c_int()
        check(lib().rb_bvh_build_cached(scene.handle, C.byref(bp), str(cache_dir).encode(), C.byref(h),
                                        C.byref(hit)))
        return cls._wrap(h.value), bool(hit.value)

    def save_binary(self, path):
        check(lib().rb_bvh_save_binary(self._h, str(path).encode()))

    def __init__(self, scene_or_arrays, platform=None, params=None):
        bp = self._params(platform, params)
        h = C.c_void_p()
        if isinstance(scene_or_arrays, Scene):
            check(lib().rb_bvh_build_host(scene_or_arrays.handle, C.byref(bp), C.byref(h)))
        else:
            v, i = scene_or_arrays
            v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
            i = np.ascontiguousarray(i, np.uint32).reshape(-1, 3)
            check(lib().rb_bvh_build_arrays(ptr(v), len(v), ptr(i), len(i), C.byref(bp), C.byref(h)))
        self._h = h.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().rb_bvh_host_release(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def stats(self):
        s = L.rb_bvh_stats()
        check(lib().rb_bvh_host_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in L.rb_bvh_stats._fields_}

    def export(self, layout):
        nb, tb, ib = C.c_size_t(), C.c_size_t(), C.c_size_t()
        check(lib().rb_bvh_export(self._h, layout, None, C.byref(nb), None, C.byref(tb), None, C.byref(ib)))
        n = np.empty(nb.value // 4, np.uint32)
        t = np.empty(tb.value // 4, np.uint32)
        i = np.empty(ib.value // 4, np.uint32)
        check(lib().rb_bvh_export(self._h, layout, ptr(n), C.byref(nb), ptr(t), C.byref(tb), ptr(i), C.byref(ib)))
        return n, t, i

    def save_text(self, path):
        check(lib().rb_bvh_save_text(self._h, path.encode()))


class CudaBvh:
    """Layout export (CudaBvh(const BVH&, BVHLayout), src/accelerators/sbvh/cudabvh.h:80-97)."""

    def __init__(self, bvh, layout=L.RB_LAYOUT_REF_COMPACT2):
        self.layout = layout
        n, t, i = bvh.export(layout)
        self.nodes = n.reshape(-1, 4)
        self.tris = t.reshape(-1, 4)
        self.index = i

    def get_node_buffer(self):
        return self.nodes

    def get_tri_woop_buffer(self):
        return self.tris

    def get_tri_index_buffer(self):
        return self.index

    @staticmethod
    def load_text(path):
        nb, tb, ib = C.c_size_t(), C.c_size_t(), C.c_size_t()
        check(lib().rb_bvh_load_text_compact2(path.encode(), None, C.byref(nb), None, C.byref(tb), None, C.byref(ib)))
        n = np.empty(nb.value // 4, np.uint32)
        t = np.empty(tb.value // 4, np.uint32)
        i = np.empty(ib.value // 4, np.uint32)
        check(lib().rb_bvh_load_text_compact2(path.encode(), ptr(n), C.byref(nb), ptr(t), C.byref(tb), ptr(i),
                                              C.byref(ib)))
        return n.reshape(-1, 4), t.reshape(-1, 4), i


class CudaTracer:
    """Device-resident BVH + scene and the batched trace call
    (CudaTracer, src/accelerators/sbvh/cudatracer.cu:43-183)."""

    def __init__(self, bvh, scene=None):
        h = C.c_void_p()
        check(lib().rb_bvh_upload(bvh.handle, scene.handle if scene is not None else None, C.byref(h)))
        self._h = h.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().rb_bvh_release(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def device_bytes(self):
        b = C.c_size_t()
        check(lib().rb_bvh_device_bytes(self._h, C.byref(b)))
        return b.value

    def trace_device(self, d_rays, n, d_hits, any_hit=False, d_mask=None, stream=None):
        flags = (L.RB_ANY if any_hit else L.RB_CLOSEST) | (L.RB_MASKED if d_mask is not None else 0)
        check(lib().rb_trace(self._h, _dp(d_rays), n, flags, _dp(d_mask), _dp(d_hits), _sp(stream)))

    def trace(self, rays, any_hit=False, mask=None, stream=None):
        """Host convenience: upload rays (n x 8 float32 or RAY_DTYPE), trace, download (ids, t)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        if n == 0:
            return np.empty(0, np.int32), np.empty(0, np.float32)
        dr = DeviceBuffer.from_array(rays, _sp(stream))
        dh = DeviceBuffer(n * 8)
        dm = DeviceBuffer.from_array(np.ascontiguousarray(mask, np.uint8), _sp(stream)) if mask is not None else None
        self.trace_device(dr, n, dh, any_hit, dm, stream)
        h = dh.download(HIT_DTYPE, n, _sp(stream))
        return h["id"].copy(), h["t"].copy()

    def trace_simple(self, rays, any_hit=False):
        """One ray per thread through the device function rb_trace_one()
        (include/rb_trace_device.h); same results as trace()."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        dr = DeviceBuffer.from_array(rays)
        dh = DeviceBuffer(max(n, 1) * 8)
        check(lib().rb_trace_simple(self._h, dr.ptr, n, L.RB_ANY if any_hit else L.RB_CLOSEST, dh.ptr, None))
        h = dh.download(HIT_DTYPE, n)
        return h["id"].copy(), h["t"].copy()

    def device_view(self):
        v = L.rb_bvh_dev()
        check(lib().rb_bvh_device_view(self._h, C.byref(v)))
        return v

    def trace_counts(self, rays, any_hit=False, wide=False):
        """(ids, t, counts[n, 3]) of the binary walk, or of the wide walk
        (closest hit) with wide=True."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        dr = DeviceBuffer.from_array(rays)
        dh = DeviceBuffer(n * 8)
        dc = DeviceBuffer(n * 12)
        flags = (L.RB_ANY if any_hit else L.RB_CLOSEST) | (L.RB_WIDE_COUNTS if wide else 0)
        check(lib().rb_trace_counts(self._h, dr.ptr, n, flags, dh.ptr, dc.ptr, None))
        h = dh.download(HIT_DTYPE, n)
        c = dc.download(np.uint32, 3 * n).reshape(n, 3)
        return h["id"].copy(), h["t"].copy(), c


def _dp(b):
    if b is None:
        return None
    return b.ptr if isinstance(b, DeviceBuffer) else b


def _sp(s):
    if s is None:
        return None
    return s.handle if isinstance(s, Stream) else s


def get_trace_config():
    c = L.rb_trace_config()
    check(lib().rb_trace_get_config(C.byref(c)))
    return {k: getattr(c, k) for k, _ in L.rb_trace_config._fields_}


def set_trace_config(**kw):
    cur = get_trace_config()
    cur.update(kw)
    c = L.rb_trace_config(**cur)
    check(lib().rb_trace_set_config(C.byref(c)))


def autotune(tracer, d_rays, n, d_hits, any_hit=False, reps=3, stream=None):
    """Pick and install the fastest traversal configuration for this batch."""
    c = L.rb_trace_config()
    ms = C.c_float()
    check(lib().rb_trace_autotune(tracer.handle, _dp(d_rays), n, L.RB_ANY if any_hit else L.RB_CLOSEST,
                                  _dp(d_hits), reps, C.byref(c), C.byref(ms), _sp(stream)))
    return {k: getattr(c, k) for k, _ in L.rb_trace_config._fields_}, ms.value


def copy_kernel(d_dst, d_src, nbytes, stream=None):
    check(lib().rb_copy_kernel(_dp(d_dst), _dp(d_src), nbytes, _sp(stream)))


def selftest_rcp():
    """(inputs taking the refined reciprocal, inputs differing from the IEEE
    division) over all 2^32 single-precision inputs (rb_selftest_rcp)."""
    fast, bad = C.c_ulonglong(0), C.c_ulonglong(0)
    check(lib().rb_selftest_rcp(C.byref(fast), C.byref(bad)))
    return fast.value, bad.value


class PerspectiveCamera:
    """Pinhole camera (PerspectiveCamera, src/core/camera.h:48-71)."""

    def __init__(self, position, lookat, up, resolution, horizontal_fov):
        self.c = L.rb_camera((C.c_float * 3)(*position), (C.c_float * 3)(*lookat), (C.c_float * 3)(*up),
                             float(horizontal_fov), int(resolution[0]), int(resolution[1]))

    @property
    def width(self):
        return self.c.width

    @property
    def height(self):
        return self.c.height

    def get_total_pixels(self):
        return self.c.width * self.c.height

    def matrices(self):
        a = np.empty(16, np.float32)
        b = np.empty(16, np.float32)
        check(lib().rb_camera_matrices(C.byref(self.c), ptr(a), ptr(b)))
        return a.reshape(4, 4), b.reshape(4, 4)

    def generate_rays_device(self, d_rays, spp=1, seed=0, jitter=True, tile=None, stream=None):
        x0, y0, w, h = tile if tile else (0, 0, self.c.width, self.c.height)
        check(lib().rb_raygen_camera(C.byref(self.c), x0, y0, w, h, spp, seed, int(jitter), _dp(d_rays), _sp(stream)))
        return w * h * spp


def trace_batch(batches, stream=None):
    """One persistent launch over many (tracer, d_rays, n, d_hits, any_hit) batches."""
    arr = (L.rb_batch_desc * max(len(batches), 1))()
    for i, (tr, dr, n, dh, anyh) in enumerate(batches):
        arr[i] = L.rb_batch_desc(tr.handle, _dp(dr), _dp(dh), n, L.RB_ANY if anyh else L.RB_CLOSEST)
    check(lib().rb_trace_batch(arr, len(batches), _sp(stream)))


FILTERS = {"box": L.RB_FILTER_BOX, "triangle": L.RB_FILTER_TRIANGLE, "gaussian": L.RB_FILTER_GAUSSIAN}
# RayCastShade names of the integrator config (src/config/configuration.cpp:222-250)
SHADES = {"primitive_id": L.RB_SHADE_PRIMITIVE_ID, "material_id": L.RB_SHADE_MATERIAL_ID,
          "normals": L.RB_SHADE_NORMALS, "uvs": L.RB_SHADE_UVS, "uniform": L.RB_SHADE_UNIFORM}


def render_params(**kw):
    """rb_render_params with the library defaults, overridden by keywords
    (filter may be a name: "box", "triangle", "gaussian"; shade one of SHADES)."""
    p = L.rb_render_params()
    lib().rb_render_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "filter" and isinstance(v, str):
            v = FILTERS[v]
        if k == "shade" and isinstance(v, str):
            v = SHADES[v]
        if k == "filter_width":
            v = (C.c_float * 2)(*v)
        setattr(p, k, v)
    return p


class SamplerRenderer:
    """Camera rays + integrator + film for one tile (SamplerRenderer::render,
    src/renderers/samplerrenderer.cpp:58-107; Integrator::render, src/core/integrator.h:59)."""

    INTEGRATORS = {"ao": "rb_render_ao", "pt": "rb_render_pt", "bdpt": "rb_render_bdpt",
                   "raycast": "rb_render_raycast"}

    def __init__(self, tracer, camera, integrator="ao", **params):
        if integrator not in self.INTEGRATORS:
            raise ValueError(f"unknown integrator {integrator!r}")
        self.tracer, self.camera, self.integrator = tracer, camera, integrator
        self.params = render_params(**params)

    def render_device(self, d_out, tile=None, stream=None):
        x0, y0, w, h = tile if tile else (0, 0, self.camera.width, self.camera.height)
        fn = getattr(lib(), self.INTEGRATORS[self.integrator])
        check(fn(self.tracer.handle, C.byref(self.camera.c), x0, y0, w, h, C.byref(self.params), _dp(d_out),
                 _sp(stream)))
        return w, h

    def render(self, tile=None, stream=None):
        """Host convenience: returns float32[h, w, 3]."""
        x0, y0, w, h = tile if tile else (0, 0, self.camera.width, self.camera.height)
        d = DeviceBuffer(w * h * 12)
        self.render_device(d, (x0, y0, w, h), stream)
        if stream is not None:
            _sync(stream)
        return d.download(np.float32, w * h * 3).reshape(h, w, 3)


class RenderGraph:
    """A render pass captured into a HIP graph over a fixed tile and output
    buffer (rb_render_graph_*): launch(seed) replays the whole pass --
    raygen, every bounce, film -- for a new seed with one graph launch."""

    KINDS = {"ao": L.RB_REQ_AO, "pt": L.RB_REQ_PT, "bdpt": L.RB_REQ_BDPT, "raycast": L.RB_REQ_RAYCAST}

    def __init__(self, tracer, camera, integrator, d_out, tile=None, **params):
        x0, y0, w, h = tile if tile else (0, 0, camera.width, camera.height)
        self.tile = (x0, y0, w, h)
        self.params = render_params(**params)
        self._keep = (tracer, camera, d_out)
        g = C.c_void_p()
        check(lib().rb_render_graph_create(self.KINDS[integrator], tracer.handle, C.byref(camera.c), x0, y0, w, h,
                                           C.byref(self.params), _dp(d_out), C.byref(g)))
        self.handle = g.value

    def launch(self, seed, stream=None):
        check(lib().rb_render_graph_launch(self.handle, seed, _sp(stream)))

    def close(self):
        if getattr(self, "handle", None):
            lib().rb_render_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_image(path, img):
    """Host float32 image [h, w, 3] to .exr / .pfm / .ppm / .png (rb_image_write)."""
    a = np.ascontiguousarray(img, np.float32)
    h, w = a.shape[:2]
    check(lib().rb_image_write(str(path).encode(), ptr(a), w, h))


def _sync(stream):
    check(lib().rb_stream_synchronize(_sp(stream)))


def raygen_diffuse_device(tracer, d_prim, d_hits, n_primary, n_out, seed, d_out, stream=None):
    check(lib().rb_raygen_diffuse(tracer.handle, _dp(d_prim), _dp(d_hits), n_primary, n_out, seed, _dp(d_out),
                                  _sp(stream)))


class RenderServer:
    """Continuous-batching request server over resident scenes (rb_server_*).

    The reference renders one configured frame per process run
    (src/core/application.cpp:225-239); this keeps scenes resident and admits
    AO / PT / BDPT tiles and raw trace batches as they arrive within a
    device-memory budget, each on its own HIP stream (mode "streams") or all
    in lockstep with one batched trace launch per round (mode "lockstep").
    Results are independent of what else runs and of the mode: every random
    number is keyed by (request seed, pixel, sample, bounce)."""

    KINDS = {"ao": L.RB_REQ_AO, "pt": L.RB_REQ_PT, "bdpt": L.RB_REQ_BDPT, "raycast": L.RB_REQ_RAYCAST}

    def __init__(self, max_inflight=8, hbm_budget=0, max_queue=0, policy="deadline", mode="streams"):
        pol = {"deadline": L.RB_SCHED_DEADLINE, "fifo": L.RB_SCHED_FIFO}[policy]
        md = {"streams": L.RB_SERVE_STREAMS, "lockstep": L.RB_SERVE_LOCKSTEP}[mode]
        cfg = L.rb_server_config(max_inflight, hbm_budget, max_queue, pol, md)
        h = C.c_void_p()
        check(lib().rb_server_create(C.byref(cfg), C.byref(h)))
        self.handle = h.value
        self._tracers = {}
        self._live = {}  # request id -> (rb_request, host image or None)

    def add_scene(self, tracer):
        sid = C.c_uint32()
        check(lib().rb_server_add_scene(self.handle, tracer.handle, C.byref(sid)))
        self._tracers[sid.value] = tracer
        return sid.value

    def remove_scene(self, sid):
        check(lib().rb_server_remove_scene(self.handle, sid))
        self._tracers.pop(sid, None)

    @staticmethod
    def render_request(kind, scene, camera, tile=None, **params):
        x0, y0, w, h = tile if tile else (0, 0, camera.width, camera.height)
        r = L.rb_request()
        r.kind = RenderServer.KINDS[kind]
        r.scene = scene
        r.camera = camera.c
        r.x0, r.y0, r.w, r.h = x0, y0, w, h
        r.params = render_params(**params)
        return r

    @staticmethod
    def request_bytes(req):
        b = C.c_uint64()
        check(lib().rb_server_request_bytes(C.byref(req), C.byref(b)))
        return b.value

    def submit(self, req, out=None):
        if out is not None:
            req.out_host = out.ctypes.data_as(C.c_void_p)
        rid = C.c_uint64()
        check(lib().rb_server_submit(self.handle, C.byref(req), C.byref(rid)))
        self._live[rid.value] = (req, out)
        return rid.value

    def submit_render(self, kind, scene, camera, tile=None, **params):
        """Queues a tile render; wait() returns its float32[h, w, 3] image."""
        req = self.render_request(kind, scene, camera, tile, **params)
        out = np.zeros((req.h, req.w, 3), np.float32)
        return self.submit(req, out)

    def submit_trace(self, scene, d_rays, d_hits, n, flags=L.RB_CLOSEST):
        req = L.rb_request()
        req.kind = L.RB_REQ_TRACE
        req.scene = scene
        req.rays, req.hits, req.n, req.flags = _dp(d_rays), _dp(d_hits), n, flags
        return self.submit(req)

    def poll(self, rid):
        st = L.rb_request_status()
        s = lib().rb_server_poll(self.handle, rid, C.byref(st))
        if s not in (L.RB_OK, L.RB_ERR_NOT_READY):
            check(s)
        return st

    def wait(self, rid):
        """Blocks until the request ends; returns (rb_request_status, image or None)."""
        st = L.rb_request_status()
        check(lib().rb_server_wait(self.handle, rid, C.byref(st)))
        _, out = self._live.pop(rid, (None, None))
        return st, out

    def drain(self):
        check(lib().rb_server_drain(self.handle))

    def stats(self):
        st = L.rb_server_stats()
        check(lib().rb_server_get_stats(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in L.rb_server_stats._fields_}

    def reset_stats(self):
        check(lib().rb_server_reset_stats(self.handle))

    def inject_alloc_failures(self, k):
        check(lib().rb_server_inject_alloc_failures(self.handle, k))

    def close(self):
        if self.handle:
            lib().rb_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
