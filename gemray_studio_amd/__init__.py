"""gemray_studio_amd: MI355X (gfx950) spectral transport for gemray's render backend.

The compute path is ``libgemray_hip.so`` (HIP kernels behind the C ABI of
``include/gemray_hip.h``); ``backend.GpuBackend`` is the host-side mirror of gemray's
``renderer::gpu_backend::GpuBackend``.
"""
from .backend import GpuBackend, HdrMap, SceneRef, Studio, camera_new  # noqa: F401
from .materials import GemMaterial, by_name  # noqa: F401
