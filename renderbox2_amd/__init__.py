"""renderbox2_amd -- MI355X-native ray-tracing core for renderbox2's hot path.

librb.so (host C++ SBVH builder + gfx950 HIP kernels) behind the C-ABI in
include/rb.h; this package mirrors the reference's Scene / BVH / CudaBvh /
CudaTracer / PerspectiveCamera interface over it.
"""
from ._lib import RbError, lib  # noqa: F401
from .core import (BVH, FLT_MAX, HIT_DTYPE, RAY_DTYPE, BuildParams, CudaBvh, CudaTracer,  # noqa: F401
                   DeviceBuffer, Event, PerspectiveCamera, Platform, RenderGraph, RenderServer, Scene, Stream, autotune, copy_kernel,
                   device_init,
                   SamplerRenderer, get_trace_config, raygen_diffuse_device, render_params, selftest_rcp, set_trace_config,
                   trace_batch, write_image)
from . import config, scenes  # noqa: F401
