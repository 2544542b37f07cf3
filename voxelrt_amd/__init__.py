"""voxelrt_amd - MI355X-native brickmap ray traversal (VoxelRT's hot path).

The product is libvoxelrt_hip.so (C ABI in include/vrt.h); this package holds
its sources (csrc/), the build script and a thin Python mirror of the
reference's voxel-map / ray-cast interface used by tests and benchmarks.
"""
from ._lib import VrtError, CapacityError, lib, LIB_PATH, VRT_ORDER_FORWARD, VRT_ORDER_REVERSE  # noqa: F401
from .api import (  # noqa: F401
    Scene,
    World,
    RayBatch,
    HitBatch,
    HitResult,
    CameraSpec,
    incoherent_rays_host,
    new_counters,
    release_stream,
    bounce,
    cast_frame,
    render_frame,
    GBufferImages,
    Batch,
    BatchPlan,
    autotune,
    render,
    render_saved,
    render_backward,
    sun_direction,
)

__all__ = [
    "VrtError", "CapacityError", "lib", "LIB_PATH", "Scene", "World", "RayBatch", "HitBatch", "HitResult",
    "CameraSpec", "incoherent_rays_host", "new_counters", "release_stream", "bounce", "cast_frame", "render_frame",
    "GBufferImages", "Batch", "BatchPlan", "autotune", "render",
    "render_saved", "render_backward", "sun_direction",
]
