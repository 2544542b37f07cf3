"""ctypes bindings of libvoxelrt_hip.so (include/vrt.h, include/vrt_scene.h).

The library is loaded from the package directory; a missing library is an
error (there is no fallback path).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libvoxelrt_hip.so"

VRT_OK = 0
VRT_ERR_INVALID_ARGUMENT = 1
VRT_ERR_CAPACITY = 2
VRT_ERR_HIP = 3
VRT_ERR_NO_DEVICE = 4
VRT_ERR_STATE = 5

VRT_SCHED_AUTO = 0
VRT_SCHED_GRID = 1
VRT_SCHED_PERSISTENT = 2
VRT_SCHED_COMPACT = 3

VRT_BOUNCE_PRIMARY_ORDER = 0
VRT_BOUNCE_DIRECTION_ORDER = 1

VRT_ORDER_FORWARD = 0
VRT_ORDER_REVERSE = 1

VRT_SEG_CAMERA = 0
VRT_SEG_RAYS = 1

_STATUS_NAMES = {
    VRT_ERR_INVALID_ARGUMENT: "invalid argument",
    VRT_ERR_CAPACITY: "capacity",
    VRT_ERR_HIP: "HIP error",
    VRT_ERR_NO_DEVICE: "no device",
    VRT_ERR_STATE: "state",
}


class VrtError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{_STATUS_NAMES.get(status, status)}] {message}")
        self.status = status


class CapacityError(VrtError):
    """Raised when brick slots are exhausted (BrickSlotAllocator.cpp:13-16)."""


P = C.c_void_p
u32 = C.c_uint32
i32 = C.c_int32
u64 = C.c_uint64
f32 = C.c_float


class WorldDesc(C.Structure):
    _fields_ = [("log2_size_xz", u32), ("log2_size_y", u32), ("initial_slots", u32), ("max_slots", u32)]


class WorldInfo(C.Structure):
    _fields_ = [("size_xz", u32), ("size_y", u32), ("sectors_xz", u32), ("sectors_y", u32),
                ("num_sectors", u32), ("num_bricks", u32), ("slot_capacity", u32), ("device_bytes", u64),
                ("slot_high_water", u32), ("generation", u32), ("upload_bytes", u64),
                ("upload_brick_bytes", u64), ("moved_bricks", u64)]


class WorldBuffers(C.Structure):
    _fields_ = [("sectors", P), ("occupancy", P), ("voxels", P), ("palette", P)]


class Rays(C.Structure):
    _fields_ = [("ox", P), ("oy", P), ("oz", P), ("dx", P), ("dy", P), ("dz", P),
                ("world_origin", i32 * 3), ("n", u32)]


class Hit(C.Structure):
    _fields_ = [("x", i32), ("y", i32), ("z", i32), ("info", u32)]


class Hits(C.Structure):
    _fields_ = [("t", P), ("rec", P), ("pos", P), ("uv", P), ("material_data", P)]


class CastCounters(C.Structure):
    _fields_ = [("rays", u64), ("iters", u64), ("descents", u64), ("hits", u64)]


class CastOpts(C.Structure):
    _fields_ = [("schedule", u32), ("refill", u32), ("chunk", u32), ("waves_per_cu", u32),
                ("counters", P), ("count", P), ("order", u32), ("pass_cap", u32)]


class HitF64(C.Structure):
    _fields_ = [("distance", C.c_double), ("normal", f32 * 3), ("uv", f32 * 2), ("voxel", i32 * 3)]


class RayBuffers(C.Structure):
    _fields_ = [("ox", P), ("oy", P), ("oz", P), ("dx", P), ("dy", P), ("dz", P)]


VRT_SHADER_FULL = 0
VRT_SHADER_COARSE = 1


class ShaderOpts(C.Structure):
    _fields_ = [("method", u32), ("ray_cell_lut", u32), ("max_iters", u32), ("reserved", u32)]


class CvoxInfo(C.Structure):
    _fields_ = [("sectors_in_file", u32), ("sectors_loaded", u32), ("sectors_outside", u32),
                ("bricks_loaded", u64)]


class RenderOpts(C.Structure):
    _fields_ = [("max_bounces", u32), ("frame", u32), ("ray_cell_lut", u32), ("order", u32), ("sky", f32 * 3),
                ("near_z", f32), ("far_z", f32)]


class GBuffer(C.Structure):
    _fields_ = [("albedo_normal", P), ("irradiance", P), ("depth", P)]


class FrameOpts(C.Structure):
    _fields_ = [("schedule", u32), ("refill", u32), ("chunk", u32), ("counters", P), ("order", u32)]


VRT_HIT_NO_RAY = 1 << 14


class Camera(C.Structure):
    _fields_ = [("world_origin", i32 * 3), ("eye", f32 * 3), ("forward", f32 * 3), ("right", f32 * 3),
                ("up", f32 * 3), ("width", u32), ("height", u32), ("tile_w", u32), ("tile_h", u32)]


class Segment(C.Structure):
    _fields_ = [("kind", u32), ("camera", u32), ("ray_offset", u32), ("hit_offset", u32), ("count", u32),
                ("world_origin", i32 * 3), ("count_ptr", P)]


class Batch(C.Structure):
    _fields_ = [("cameras", P), ("ncameras", u32), ("segments", P), ("nsegments", u32),
                ("rays", Rays), ("hits", Hits), ("max_iters", u32)]


class BatchTuning(C.Structure):
    _fields_ = [("tile_w", u32), ("tile_h", u32), ("rays_per_lane", u32), ("chunk", u32)]


class SectorStream(C.Structure):
    _fields_ = [("log2_size_xz", u32), ("log2_size_y", u32), ("nsec", u32), ("sector_idx", P),
                ("alloc_mask", P), ("bricks", P), ("nbricks", u64), ("solid_voxels", u64)]


class TerrainParams(C.Structure):
    _fields_ = [("log2_size_xz", u32), ("log2_size_y", u32), ("seed", u64), ("octaves", u32),
                ("frequency", f32), ("gain", f32), ("base", f32), ("amp", f32), ("cave_threshold", f32),
                ("cave_octaves", u32), ("cave_frequency", f32), ("band", u32)]


# name -> (restype, argtypes)
_SIGNATURES = {
    "vrt_last_error": (C.c_char_p, []),
    "vrt_abi_version": (C.c_int, []),
    "vrt_world_create": (C.c_int, [C.POINTER(WorldDesc), P, C.POINTER(P)]),
    "vrt_world_destroy": (C.c_int, [P]),
    "vrt_world_get_info": (C.c_int, [P, C.POINTER(WorldInfo)]),
    "vrt_world_sync": (C.c_int, [P, u32, P, P, P, P]),
    "vrt_world_sync_bricks": (C.c_int, [P, u32, P, P, P, P, P]),
    "vrt_world_set_palette": (C.c_int, [P, P, P]),
    "vrt_world_get_buffers": (C.c_int, [P, C.POINTER(WorldBuffers)]),
    "vrt_cast": (C.c_int, [P, C.POINTER(Rays), u32, C.POINTER(Hits), P]),
    "vrt_cast_ex": (C.c_int, [P, C.POINTER(Rays), u32, C.POINTER(Hits), C.POINTER(CastOpts), P]),
    "vrt_stream_release": (C.c_int, [P]),
    "vrt_cast_one_f64": (C.c_int, [P, C.POINTER(C.c_double * 3), C.POINTER(C.c_double * 3), u32, C.POINTER(HitF64)]),
    "vrt_bounce_scratch_bytes": (C.c_size_t, [u32]),
    "vrt_bounce_generate": (C.c_int, [C.POINTER(Rays), C.POINTER(Hits), u32, u32, C.POINTER(RayBuffers), P, P, P, P]),
    "vrt_bounce_generate_ex": (C.c_int, [C.POINTER(Rays), C.POINTER(Hits), u32, u32, u32, C.POINTER(RayBuffers), P, P, P, P]),
    "vrt_bounce_generate_noalloc": (C.c_int, [C.POINTER(Rays), C.POINTER(Hits), u32, u32, C.POINTER(RayBuffers), P, P,
                                              P]),
    "vrt_camera_rays": (C.c_int, [C.POINTER(Camera), C.POINTER(RayBuffers), P]),
    "vrt_cast_shader": (C.c_int, [P, C.POINTER(Rays), C.POINTER(ShaderOpts), C.POINTER(Hits), P]),
    "vrt_render_frame": (C.c_int, [P, C.POINTER(Camera), C.POINTER(RenderOpts), C.POINTER(GBuffer), P]),
    "vrt_cast_frame": (C.c_int, [P, C.POINTER(Camera), u32, u32, C.POINTER(Hits), C.POINTER(Hits),
                                 C.POINTER(FrameOpts), P]),
    "vrt_batch_plan_create": (C.c_int, [P, C.POINTER(Batch), C.POINTER(BatchTuning), P, C.POINTER(P)]),
    "vrt_batch_launch": (C.c_int, [P, P, P]),
    "vrt_batch_graph_launch": (C.c_int, [P, P]),
    "vrt_batch_plan_destroy": (C.c_int, [P]),
    "vrt_batch_autotune": (C.c_int, [P, C.POINTER(Batch), C.c_char_p, C.POINTER(BatchTuning), C.POINTER(f32), P]),
    "vrt_render_fwd": (C.c_int, [P, C.POINTER(Camera), u32, u32, C.POINTER(f32 * 3), f32, P, P, P, P, P]),
    "vrt_render_bwd": (C.c_int, [P, u32, P, P, P, P, P]),
    "vrt_render_bwd_workspace_bytes": (C.c_size_t, [u32, u64]),
    "vrt_render_bwd_bricks": (C.c_int, [P, u32, P, P, P, u64, P, P, C.c_size_t, P]),
    "vrt_render_bwd_sparse": (C.c_int, [P, u32, P, P, P, P, P, P]),
    # vrt_scene.h
    "vrt_scene_terrain": (C.c_int, [C.POINTER(TerrainParams), u32, C.POINTER(P)]),
    "vrt_scene_spheres": (C.c_int, [u32, u32, u32, f32, f32, u64, u32, C.POINTER(P)]),
    "vrt_scene_from_dense": (C.c_int, [u32, u32, P, C.POINTER(P)]),
    "vrt_scene_get": (C.c_int, [P, C.POINTER(SectorStream)]),
    "vrt_scene_voxel": (C.c_uint8, [P, i32, i32, i32]),
    "vrt_scene_set_voxel": (C.c_int, [P, i32, i32, i32, C.c_uint8]),
    "vrt_scene_set_voxels": (C.c_int, [P, u64, P, P]),
    "vrt_scene_mark_all_dirty": (C.c_int, [P]),
    "vrt_scene_delete_empty_bricks": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "vrt_scene_dirty_sectors": (u32, [P]),
    "vrt_scene_destroy": (None, [P]),
    "vrt_scene_load_cvox": (C.c_int, [C.c_char_p, u32, u32, C.POINTER(i32 * 3), C.POINTER(P), C.POINTER(CvoxInfo)]),
    "vrt_scene_save_cvox": (C.c_int, [P, C.POINTER(i32 * 3), P, C.c_char_p]),
    "vrt_scene_get_palette": (C.c_int, [P, P]),
    "vrt_scene_set_palette": (C.c_int, [P, P]),
    "vrt_world_from_scene": (C.c_int, [P, P, C.POINTER(P)]),
    "vrt_world_sync_scene": (C.c_int, [P, P, P]),
    "vrt_camera_look": (C.c_int, [C.POINTER(C.c_double * 3), C.c_double, C.c_double, C.c_double, u32, u32, u32, u32,
                                  C.POINTER(Camera)]),
    "vrt_camera_rays_host": (C.c_int, [C.POINTER(Camera), C.POINTER(P * 6)]),
    "vrt_incoherent_rays_host": (C.c_int, [u32, u32, u32, u64, C.POINTER(P * 6)]),
    "vrt_bounce_host": (C.c_int, [u32, P, P, u32, u32, C.POINTER(P * 6), P, P]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libvoxelrt_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m voxelrt_amd.build`")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(status: int) -> None:
    if status != VRT_OK:
        msg = lib().vrt_last_error().decode(errors="replace")
        if status == VRT_ERR_CAPACITY:
            raise CapacityError(status, msg)
        raise VrtError(status, msg)
