"""ctypes binding of librb.so (the C-ABI declared in include/rb.h)."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# RB_LIB selects another build of the same library (the sanitizer build of
# `make -C renderbox2_amd/csrc -f asan.mk asan` is librb_asan.so).
LIB_PATH = os.environ.get("RB_LIB") or os.path.join(_HERE, "librb.so")

P = C.c_void_p
U32 = C.c_uint32
U64 = C.c_uint64
F32 = C.c_float
I32 = C.c_int32
SZ = C.c_size_t

RB_OK, RB_ERR_INVALID, RB_ERR_OOM, RB_ERR_DEVICE, RB_ERR_BUSY, RB_ERR_UNSUPPORTED, RB_ERR_NOT_READY = range(7)
RB_CLOSEST, RB_ANY, RB_MASKED, RB_WIDE_COUNTS = 0, 1, 2, 4
RB_LAYOUT_REF_COMPACT2, RB_LAYOUT_MI355X, RB_LAYOUT_MI355X_WIDE = 1, 2, 3
(RB_MAT_LAMBERT, RB_MAT_MIRROR, RB_MAT_EMITTER, RB_MAT_GLASS, RB_MAT_ORENNAYAR, RB_MAT_MICROFACET,
 RB_MAT_LAYERED) = range(7)
RB_FRESNEL_NONE, RB_FRESNEL_CONDUCTOR, RB_FRESNEL_DIELECTRIC = 0, 1, 2
RB_GEN_ICOSPHERES, RB_GEN_THIN_QUADS, RB_GEN_ROOM, RB_GEN_EMITTER_QUADS = 0, 1, 2, 3


class RbError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


_STATUS = {0: "RB_OK", 1: "RB_ERR_INVALID", 2: "RB_ERR_OOM", 3: "RB_ERR_DEVICE", 4: "RB_ERR_BUSY",
           5: "RB_ERR_UNSUPPORTED", 6: "RB_ERR_NOT_READY"}


class rb_material(C.Structure):
    _fields_ = [("kind", U32), ("color", F32 * 3), ("param", F32 * 8)]


class rb_gen_params(C.Structure):
    _fields_ = [("kind", U32), ("count", U32), ("subdiv", U32), ("seed", U64),
                ("r0", F32), ("r1", F32), ("c0", F32), ("c1", F32),
                ("material_id", I32), ("mix_materials", U32), ("emit", F32 * 3),
                ("subdiv_hi", U32), ("hi_count", U32), ("smooth", U32)]


class rb_build_params(C.Structure):
    _fields_ = [("min_leaf", U32), ("max_leaf", U32), ("split_alpha", F32), ("node_cost", F32),
                ("tri_cost", F32), ("threads", U32)]


class rb_bvh_stats(C.Structure):
    _fields_ = [("sah", F32), ("n_inner", U32), ("n_leaf", U32), ("n_tri_refs", U32),
                ("n_child_nodes", U32), ("max_depth", U32), ("duplicates", U64), ("build_ms", C.c_double)]


class rb_trace_config(C.Structure):
    _fields_ = [("block", U32), ("lds_stack", U32), ("chunk", U32), ("fetch_threshold", U32),
                ("waves_per_cu", U32), ("speculative", U32), ("leaf_gate", U32), ("lds_nodes", U32),
                ("wide", U32)]


class rb_camera(C.Structure):
    _fields_ = [("position", F32 * 3), ("lookat", F32 * 3), ("up", F32 * 3), ("hfov_deg", F32),
                ("width", U32), ("height", U32)]


class rb_isect_soa(C.Structure):
    _fields_ = [("position", P * 3), ("geometric_normal", P * 3), ("shading_normal", P * 3),
                ("incoming_direction", P * 3), ("uv", P), ("primitive_id", P), ("material_id", P),
                ("front", P), ("intersected", P), ("epsilon", P), ("t", P)]


# name -> (restype, argtypes); every function returns rb_status unless stated.
_SIGS = {
    "rb_last_error": (SZ, [C.c_char_p, SZ]),
    "rb_status_string": (C.c_char_p, [C.c_int]),
    "rb_version": (C.c_char_p, []),
    "rb_device_init": (C.c_int, [C.c_int]),
    "rb_device_name": (C.c_int, [C.c_char_p, SZ, P]),
    "rb_malloc": (C.c_int, [P, SZ]),
    "rb_free": (C.c_int, [P]),
    "rb_memcpy_htod": (C.c_int, [P, P, SZ, P]),
    "rb_memcpy_dtoh": (C.c_int, [P, P, SZ, P]),
    "rb_memcpy_dtod": (C.c_int, [P, P, SZ, P]),
    "rb_memset": (C.c_int, [P, C.c_int, SZ, P]),
    "rb_mem_info": (C.c_int, [P, P, P]),
    "rb_stream_create": (C.c_int, [P]),
    "rb_stream_destroy": (C.c_int, [P]),
    "rb_stream_synchronize": (C.c_int, [P]),
    "rb_event_create": (C.c_int, [P]),
    "rb_event_destroy": (C.c_int, [P]),
    "rb_event_record": (C.c_int, [P, P]),
    "rb_event_elapsed_ms": (C.c_int, [P, P, P]),
    "rb_event_synchronize": (C.c_int, [P]),
    "rb_scene_create": (C.c_int, [P]),
    "rb_scene_destroy": (C.c_int, [P]),
    "rb_scene_add_material": (C.c_int, [P, P, P]),
    "rb_scene_add_layered_material": (C.c_int, [P, P, U32, P]),
    "rb_scene_add_mesh": (C.c_int, [P, P, U32, P, U32, U32]),
    "rb_scene_add_mesh_ex": (C.c_int, [P, P, U32, P, U32, P, U32, P, P, U32, P, U32]),
    "rb_scene_shading_counts": (C.c_int, [P, P, P, P, P]),
    "rb_scene_shading": (C.c_int, [P, P, P, P, P]),
    "rb_scene_counts": (C.c_int, [P, P, P, P, P]),
    "rb_scene_geometry": (C.c_int, [P, P, P, P, P]),
    "rb_scene_materials": (C.c_int, [P, P]),
    "rb_scene_bounds": (C.c_int, [P, P, P]),
    "rb_scene_generate": (C.c_int, [P, P]),
    "rb_scene_load_obj": (C.c_int, [P, C.c_char_p, U32]),
    "rb_scene_load_xml": (C.c_int, [P, C.c_char_p, U32, U32, P]),
    "rb_image_write": (C.c_int, [C.c_char_p, P, U32, U32]),
    "rb_build_params_default": (None, [P]),
    "rb_bvh_build_host": (C.c_int, [P, P, P]),
    "rb_bvh_build_arrays": (C.c_int, [P, U32, P, U32, P, P]),
    "rb_bvh_host_stats": (C.c_int, [P, P]),
    "rb_bvh_host_release": (C.c_int, [P]),
    "rb_bvh_export": (C.c_int, [P, U32, P, P, P, P, P, P]),
    "rb_bvh_save_text": (C.c_int, [P, C.c_char_p]),
    "rb_bvh_load_text_compact2": (C.c_int, [C.c_char_p, P, P, P, P, P, P]),
    "rb_bvh_save_binary": (C.c_int, [P, C.c_char_p]),
    "rb_bvh_load_binary": (C.c_int, [C.c_char_p, P, P, P]),
    "rb_bvh_build_cached": (C.c_int, [P, P, C.c_char_p, P, P]),
    "rb_bvh_upload": (C.c_int, [P, P, P]),
    "rb_bvh_release": (C.c_int, [P]),
    "rb_bvh_device_bytes": (C.c_int, [P, P]),
    "rb_trace": (C.c_int, [P, P, U32, U32, P, P, P]),
    "rb_trace_fill": (C.c_int, [P, P, U32, U32, P, P, P, P]),
    "rb_trace_fill_noalloc": (C.c_int, [P, P, U32, U32, P, P, P]),
    "rb_trace_get_config": (C.c_int, [P]),
    "rb_trace_set_config": (C.c_int, [P]),
    "rb_trace_counts": (C.c_int, [P, P, U32, U32, P, P, P]),
    "rb_trace_errors": (C.c_int, [P]),
    "rb_trace_autotune": (C.c_int, [P, P, U32, U32, P, U32, P, P, P]),
    "rb_copy_kernel": (C.c_int, [P, P, SZ, P]),
    "rb_selftest_rcp": (C.c_int, [P, P]),
    "rb_camera_matrices": (C.c_int, [P, P, P]),
    "rb_raygen_camera": (C.c_int, [P, U32, U32, U32, U32, U32, U64, C.c_int, P, P]),
    "rb_raygen_diffuse": (C.c_int, [P, P, P, U32, U32, U64, P, P]),
    "rb_trace_batch": (C.c_int, [P, U32, P]),
    "rb_bvh_device_view": (C.c_int, [P, P]),
    "rb_trace_simple": (C.c_int, [P, P, U32, U32, P, P]),
    "rb_trace_simple_fill": (C.c_int, [P, P, U32, U32, P, P]),
    "rb_render_params_default": (None, [P]),
    "rb_render_ao": (C.c_int, [P, P, U32, U32, U32, U32, P, P, P]),
    "rb_render_pt": (C.c_int, [P, P, U32, U32, U32, U32, P, P, P]),
    "rb_render_bdpt": (C.c_int, [P, P, U32, U32, U32, U32, P, P, P]),
    "rb_render_raycast": (C.c_int, [P, P, U32, U32, U32, U32, P, P, P]),
    "rb_render_graph_create": (C.c_int, [U32, P, P, U32, U32, U32, U32, P, P, P]),
    "rb_render_graph_launch": (C.c_int, [P, U64, P]),
    "rb_render_graph_destroy": (C.c_int, [P]),
    "rb_server_create": (C.c_int, [P, P]),
    "rb_server_destroy": (C.c_int, [P]),
    "rb_server_add_scene": (C.c_int, [P, P, P]),
    "rb_server_remove_scene": (C.c_int, [P, U32]),
    "rb_server_submit": (C.c_int, [P, P, P]),
    "rb_server_poll": (C.c_int, [P, U64, P]),
    "rb_server_wait": (C.c_int, [P, U64, P]),
    "rb_server_drain": (C.c_int, [P]),
    "rb_server_get_stats": (C.c_int, [P, P]),
    "rb_server_reset_stats": (C.c_int, [P]),
    "rb_server_request_bytes": (C.c_int, [P, P]),
    "rb_server_inject_alloc_failures": (C.c_int, [P, U32]),
}


RB_REQ_AO, RB_REQ_PT, RB_REQ_BDPT, RB_REQ_TRACE, RB_REQ_RAYCAST = 0, 1, 2, 3, 4
RB_SHADE_PRIMITIVE_ID, RB_SHADE_MATERIAL_ID, RB_SHADE_NORMALS, RB_SHADE_UVS, RB_SHADE_UNIFORM = 0, 1, 2, 3, 4
RB_FILTER_BOX, RB_FILTER_TRIANGLE, RB_FILTER_GAUSSIAN = 0, 1, 2
RB_REQ_QUEUED, RB_REQ_RUNNING, RB_REQ_DONE, RB_REQ_FAILED = 0, 1, 2, 3


RB_SCHED_DEADLINE, RB_SCHED_FIFO = 0, 1
RB_SERVE_STREAMS, RB_SERVE_LOCKSTEP = 0, 1


class rb_server_config(C.Structure):
    _fields_ = [("max_inflight", U32), ("hbm_budget", U64), ("max_queue", U32), ("policy", U32), ("mode", U32)]


class rb_request_status(C.Structure):
    _fields_ = [("state", C.c_int32), ("result", C.c_int32), ("queue_ms", C.c_double), ("latency_ms", C.c_double),
                ("rays", U64)]


class rb_server_stats(C.Structure):
    _fields_ = [("submitted", U64), ("completed", U64), ("failed", U64), ("rejected", U64), ("deferrals", U64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("rays", U64), ("first_submit_to_last_done_s", C.c_double), ("peak_budget_bytes", U64),
                ("peak_arena_bytes", U64), ("arena_misses", U64)]


class rb_bvh_dev(C.Structure):
    _fields_ = [("nodes", P), ("tris", P), ("index", P), ("n_nodes", U32), ("n_refs", U32), ("errors", P),
                ("wnodes", P), ("n_wnodes", U32), ("n_tris", U32), ("tri_v", P), ("tri_n", P)]


class rb_batch_desc(C.Structure):
    _fields_ = [("bvh", P), ("rays", P), ("hits", P), ("n", U32), ("flags", U32)]


class rb_render_params(C.Structure):
    _fields_ = [("spp", U32), ("seed", U64), ("ao_samples", U32), ("ao_radius", F32), ("max_depth", U32),
                ("rr_start", U32), ("rr_prob", F32), ("light_paths", U32), ("lvc_capacity", U32),
                ("connections", U32), ("light_max_depth", U32), ("ray_counter", P), ("filter", U32),
                ("filter_width", F32 * 2), ("filter_alpha", F32),
                ("shade", U32), ("path_counter", P), ("light_path_counter", P)]


class rb_request(C.Structure):
    _fields_ = [("kind", U32), ("scene", U32), ("camera", rb_camera), ("x0", U32), ("y0", U32), ("w", U32),
                ("h", U32), ("params", rb_render_params), ("out_host", P), ("rays", P), ("hits", P), ("n", U32),
                ("flags", U32)]

_L = None


def lib():
    """Load librb.so.  Raises if it has not been built: there is no fallback."""
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C renderbox2_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _L = L
    return _L


def last_error():
    buf = C.create_string_buffer(2048)
    lib().rb_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status):
    if status != RB_OK:
        raise RbError(status, last_error())
    return status


def ptr(a):
    """Host pointer of a numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(P)


def exported_symbols():
    return list(_SIGS)
