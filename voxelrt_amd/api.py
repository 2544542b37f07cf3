"""Python mirror of the reference's voxel-map and ray-cast interface.

    reference (src/VoxelRT)                        here
    VoxelMap + TerrainGenerator (host world)       Scene (seeded generators, sector stream)
    GpuVoxelStorage::SyncBuffers (device world)    World.sync / World.from_scene
    RayCast(storage, origin, dir, mask, W)         World.cast(RayBatch) -> HitBatch
    HitResult VoxelMap::RayCast(dvec3, dvec3, n)   World.ray_cast(origin, dir, max_iters=1024)

Device memory is held in torch tensors (plumbing only); all traversal work is
done by libvoxelrt_hip.so through its C ABI.  Every failure raises VrtError.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L


def _stream_handle(stream) -> int | None:
    if stream is None:
        # torch's current stream; the null stream when torch sees no device (the library reports the error)
        return torch.cuda.current_stream().cuda_stream if torch.cuda.is_available() else 0
    if isinstance(stream, torch.cuda.Stream):
        return stream.cuda_stream
    return int(stream)


def _ptr(t) -> int:
    if t is None:
        return 0
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


# ------------------------------------------------------------------ scenes


class Scene:
    """Host world as a sector stream (vrt_scene)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        self._edited = False  # once edited, stream() copies (its arrays are re-packed on edits)

    @staticmethod
    def terrain(log2_xz: int, log2_y: int, seed: int, *, octaves=4, frequency=1 / 64, gain=0.5, base=32.0,
                amp=128.0, cave_threshold=0.0, cave_octaves=3, cave_frequency=1 / 32, band=16,
                threads=0) -> "Scene":
        p = L.TerrainParams(log2_xz, log2_y, seed, octaves, frequency, gain, base, amp, cave_threshold,
                            cave_octaves, cave_frequency, band)
        out = C.c_void_p()
        L.check(L.lib().vrt_scene_terrain(C.byref(p), threads, C.byref(out)))
        return Scene(out.value)

    @staticmethod
    def spheres(log2_xz: int, log2_y: int, count: int, rmin: float, rmax: float, seed: int, threads=0) -> "Scene":
        out = C.c_void_p()
        L.check(L.lib().vrt_scene_spheres(log2_xz, log2_y, count, rmin, rmax, seed, threads, C.byref(out)))
        return Scene(out.value)

    @staticmethod
    def from_dense(voxels: np.ndarray) -> "Scene":
        """voxels: uint8 [Y][Z][X], powers of two >= 32."""
        v = np.ascontiguousarray(voxels, dtype=np.uint8)
        sy, sz, sx = v.shape
        assert sx == sz, "x and z extents must match"
        out = C.c_void_p()
        L.check(L.lib().vrt_scene_from_dense(int(sx).bit_length() - 1, int(sy).bit_length() - 1,
                                             v.ctypes.data, C.byref(out)))
        return Scene(out.value)

    @staticmethod
    def load_cvox(path: str, log2_xz: int, log2_y: int, sector_offset=(0, 0, 0)) -> tuple["Scene", dict]:
        """VoxelMap::Deserialize (VoxelMap.cpp:211-244): world sector s lands at scene sector
        s - sector_offset.  Returns (scene, info dict)."""
        out = C.c_void_p()
        info = L.CvoxInfo()
        L.check(L.lib().vrt_scene_load_cvox(str(path).encode(), log2_xz, log2_y,
                                            C.byref((C.c_int32 * 3)(*sector_offset)), C.byref(out), C.byref(info)))
        return Scene(out.value), {k: getattr(info, k) for k, _ in L.CvoxInfo._fields_}

    def save_cvox(self, path: str, sector_offset=(0, 0, 0), materials: np.ndarray | None = None) -> None:
        """VoxelMap::Serialize (VoxelMap.cpp:245-274); materials: 256 x 16-byte Material records."""
        m = None
        if materials is not None:
            m = np.ascontiguousarray(materials, dtype=np.uint8).reshape(-1)
            assert m.size == 4096
        L.check(L.lib().vrt_scene_save_cvox(self._h, C.byref((C.c_int32 * 3)(*sector_offset)),
                                            m.ctypes.data if m is not None else None, str(path).encode()))

    def palette(self) -> np.ndarray:
        """The scene's encoded palette (Material::GetEncoded per entry): a loaded file's or one
        set with set_palette."""
        pal = np.zeros(256, dtype=np.uint64)
        L.check(L.lib().vrt_scene_get_palette(self._h, pal.ctypes.data))
        return pal

    def set_palette(self, palette: np.ndarray) -> None:
        """VoxelMap::Palette: uploaded by every World.sync_scene."""
        p = np.ascontiguousarray(palette, dtype=np.uint64)
        assert p.shape == (256,)
        L.check(L.lib().vrt_scene_set_palette(self._h, p.ctypes.data))

    def stream(self) -> dict:
        """The sector stream: views into the scene's memory (the dict keeps the scene alive), or
        copies once the scene has been edited, since the library re-packs the stream's arrays
        when it folds edits in and earlier views would then point at freed memory."""
        s = L.SectorStream()
        L.check(L.lib().vrt_scene_get(self._h, C.byref(s)))
        copy = self._edited

        def view(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            buf = (C.c_uint8 * (n * np.dtype(dt).itemsize)).from_address(ptr)
            a = np.frombuffer(buf, dtype=dt)
            return a.copy() if copy else a

        return {
            "log2_size_xz": s.log2_size_xz,
            "log2_size_y": s.log2_size_y,
            "nsec": s.nsec,
            "sector_idx": view(s.sector_idx, s.nsec, np.uint32),
            "alloc_mask": view(s.alloc_mask, s.nsec, np.uint64),
            "bricks": view(s.bricks, int(s.nbricks) * 512, np.uint8),
            "nbricks": int(s.nbricks),
            "solid_voxels": int(s.solid_voxels),
            "_owner": self,
        }

    def voxel(self, x: int, y: int, z: int) -> int:
        """VoxelMap::Get: 0 outside the scene or in an absent brick."""
        return int(L.lib().vrt_scene_voxel(self._h, x, y, z))

    def set_voxel(self, x: int, y: int, z: int, value: int) -> None:
        """VoxelMap::Set: writes one voxel, allocating its brick, and marks the brick dirty;
        positions outside the scene are ignored."""
        self._edited = True
        L.check(L.lib().vrt_scene_set_voxel(self._h, x, y, z, value))

    def set_voxels(self, xyz: np.ndarray, values) -> None:
        """Many set_voxel writes in one call, in order: xyz [n, 3] (x, y, z), values [n] or one
        value for all."""
        p = np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1, 3)
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.uint8), (p.shape[0],)))
        self._edited = True
        L.check(L.lib().vrt_scene_set_voxels(self._h, p.shape[0], p.ctypes.data, v.ctypes.data))

    def mark_all_dirty(self) -> None:
        """VoxelMap::MarkAllDirty: every sector dirty with its whole allocation mask."""
        L.check(L.lib().vrt_scene_mark_all_dirty(self._h))

    def delete_empty_bricks(self) -> int:
        """Deletes bricks without a solid voxel (marking them dirty); returns how many."""
        self._edited = True
        n = C.c_uint64()
        L.check(L.lib().vrt_scene_delete_empty_bricks(self._h, C.byref(n)))
        return int(n.value)

    @property
    def dirty_sectors(self) -> int:
        """Number of sectors in the dirty set (VoxelMap::DirtyLocs)."""
        return int(L.lib().vrt_scene_dirty_sectors(self._h))

    def close(self) -> None:
        if self._h and self._h.value:
            L.lib().vrt_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ rays and hits


@dataclass
class RayBatch:
    """n rays, SoA float32 device tensors; origins relative to world_origin."""

    o: torch.Tensor          # [3, n]
    d: torch.Tensor          # [3, n]
    world_origin: tuple

    @property
    def n(self) -> int:
        return int(self.o.shape[1])

    def to_c(self) -> L.Rays:
        o, d = self.o, self.d
        r = L.Rays(o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr(), d[0].data_ptr(), d[1].data_ptr(),
                   d[2].data_ptr(), (C.c_int32 * 3)(*self.world_origin), self.n)
        return r

    @staticmethod
    def from_numpy(o: np.ndarray, d: np.ndarray, world_origin=(0, 0, 0), device="cuda") -> "RayBatch":
        return RayBatch(torch.from_numpy(np.ascontiguousarray(o, dtype=np.float32)).to(device),
                        torch.from_numpy(np.ascontiguousarray(d, dtype=np.float32)).to(device),
                        tuple(int(x) for x in world_origin))

    @staticmethod
    def empty(n: int, world_origin=(0, 0, 0), device="cuda") -> "RayBatch":
        return RayBatch(torch.empty((3, n), dtype=torch.float32, device=device),
                        torch.empty((3, n), dtype=torch.float32, device=device),
                        tuple(int(x) for x in world_origin))

    def buffers(self) -> L.RayBuffers:
        o, d = self.o, self.d
        return L.RayBuffers(o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr(), d[0].data_ptr(), d[1].data_ptr(),
                            d[2].data_ptr())


@dataclass
class HitBatch:
    t: torch.Tensor          # [n] float32, -1 on a miss
    rec: torch.Tensor        # [n, 4] int32: x, y, z, info
    pos: torch.Tensor | None  # [3, n] float32
    uv: torch.Tensor | None = None             # [2, n] float32 (vrt_cast only)
    material_data: torch.Tensor | None = None  # [n] int32: palette entry low 32 bits (vrt_cast only)

    @staticmethod
    def empty(n: int, with_pos=False, device="cuda", with_uv=False, with_material=False) -> "HitBatch":
        return HitBatch(torch.empty(n, dtype=torch.float32, device=device),
                        torch.empty((n, 4), dtype=torch.int32, device=device),
                        torch.empty((3, n), dtype=torch.float32, device=device) if with_pos else None,
                        torch.empty((2, n), dtype=torch.float32, device=device) if with_uv else None,
                        torch.empty(n, dtype=torch.int32, device=device) if with_material else None)

    def to_c(self) -> L.Hits:
        return L.Hits(self.t.data_ptr(), self.rec.data_ptr(), _ptr(self.pos), _ptr(self.uv), _ptr(self.material_data))

    # decoding of vrt_hit.info
    def material(self) -> torch.Tensor:
        return self.rec[:, 3] & 0xFF

    def iterations(self) -> torch.Tensor:
        return (self.rec[:, 3] >> 16) & 0xFFFF

    def normal(self) -> torch.Tensor:
        info = self.rec[:, 3]
        comps = [((info >> (8 + 2 * a)) & 3) for a in range(3)]
        return torch.stack([(c ^ 2) - 2 for c in comps], dim=1)


@dataclass
class HitResult:
    """VoxelMap::HitResult (VoxelMap.h:171-178)."""

    distance: float
    normal: tuple
    uv: tuple
    voxel_pos: tuple

    def is_miss(self) -> bool:
        return self.distance <= 0.0


class CameraSpec:
    """vrt_camera built by vrt_camera_look.

    anchor: the rays' integer world origin W (eye = pos - W, rounded to f32):
      "floor"   floor(pos), the reference's choice (eye in [0, 1));
      "sector"  floor(pos) rounded down to a multiple of 32 per axis, the corner of the camera's
                sector (eye in [0, 32)): the casts take their aligned-origin step, 3 VALU
                instructions shorter;
      "zero"    0, world-space rays: the step also drops the additions of W (6 shorter).
    The rays are the same world rays up to the f32 rounding of the eye."""

    ANCHORS = ("floor", "sector", "zero")

    def __init__(self, pos, yaw_deg: float, pitch_deg: float, width: int, height: int, fov_deg: float = 90.0,
                 tile=(1, 1), anchor: str = "floor"):
        if anchor not in self.ANCHORS:
            raise ValueError(f"anchor must be one of {self.ANCHORS}")
        cam = L.Camera()
        L.check(L.lib().vrt_camera_look(C.byref((C.c_double * 3)(*pos)), yaw_deg, pitch_deg, fov_deg, width,
                                        height, tile[0], tile[1], C.byref(cam)))
        if anchor != "floor":
            for a in range(3):
                w = math.floor(pos[a]) & ~31 if anchor == "sector" else 0
                cam.world_origin[a] = w
                cam.eye[a] = float(pos[a]) - w  # exact in double; ctypes rounds to f32 like vrt_camera_look
        self.c = cam

    @property
    def n(self) -> int:
        return self.c.width * self.c.height

    @property
    def world_origin(self) -> tuple:
        return tuple(self.c.world_origin)

    def rays_device(self, device="cuda", stream=None) -> RayBatch:
        rb = RayBatch.empty(self.n, self.world_origin, device)
        bufs = rb.buffers()
        L.check(L.lib().vrt_camera_rays(C.byref(self.c), C.byref(bufs), _stream_handle(stream)))
        return rb

    def rays_host(self) -> tuple[np.ndarray, np.ndarray]:
        out = np.empty((6, self.n), dtype=np.float32)
        ptrs = (C.c_void_p * 6)(*[out[i].ctypes.data for i in range(6)])
        L.check(L.lib().vrt_camera_rays_host(C.byref(self.c), C.byref(ptrs)))
        return out[:3].copy(), out[3:].copy()


def incoherent_rays_host(n: int, size_xz: int, size_y: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    out = np.empty((6, n), dtype=np.float32)
    ptrs = (C.c_void_p * 6)(*[out[i].ctypes.data for i in range(6)])
    L.check(L.lib().vrt_incoherent_rays_host(n, size_xz, size_y, seed, C.byref(ptrs)))
    return out[:3].copy(), out[3:].copy()


# ------------------------------------------------------------------ device world


class World:
    """Device brickmap (vrt_world)."""

    def __init__(self, log2_xz: int, log2_y: int, initial_slots: int = 0, max_slots: int = 0, stream=None):
        desc = L.WorldDesc(log2_xz, log2_y, initial_slots, max_slots)
        out = C.c_void_p()
        L.check(L.lib().vrt_world_create(C.byref(desc), _stream_handle(stream), C.byref(out)))
        self._h = out

    @staticmethod
    def from_scene(scene: Scene, stream=None) -> "World":
        out = C.c_void_p()
        L.check(L.lib().vrt_world_from_scene(scene._h, _stream_handle(stream), C.byref(out)))
        w = World.__new__(World)
        w._h = out
        return w

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def sync(self, sector_idx: np.ndarray, alloc_mask: np.ndarray, bricks: np.ndarray, stream=None) -> None:
        si = np.ascontiguousarray(sector_idx, dtype=np.uint32)
        am = np.ascontiguousarray(alloc_mask, dtype=np.uint64)
        br = np.ascontiguousarray(bricks, dtype=np.uint8)
        L.check(L.lib().vrt_world_sync(self._h, len(si), si.ctypes.data, am.ctypes.data,
                                       br.ctypes.data if br.size else 0, _stream_handle(stream)))

    def sync_bricks(self, sector_idx: np.ndarray, alloc_mask: np.ndarray, dirty_mask: np.ndarray,
                    bricks: np.ndarray, stream=None) -> None:
        """DirtyLocs sync (GpuRenderer.cpp:45-167): bricks holds only the dirty bricks of each
        sector (dirty_mask & alloc_mask), sector after sector, in increasing brick index."""
        si = np.ascontiguousarray(sector_idx, dtype=np.uint32)
        am = np.ascontiguousarray(alloc_mask, dtype=np.uint64)
        dm = np.ascontiguousarray(dirty_mask, dtype=np.uint64)
        br = np.ascontiguousarray(bricks, dtype=np.uint8)
        L.check(L.lib().vrt_world_sync_bricks(self._h, len(si), si.ctypes.data, am.ctypes.data, dm.ctypes.data,
                                              br.ctypes.data if br.size else 0, _stream_handle(stream)))

    def sync_scene(self, scene: Scene, stream=None) -> None:
        """SyncBuffers(VoxelMap&): consumes the scene's dirty set (vrt_world_sync_scene)."""
        L.check(L.lib().vrt_world_sync_scene(self._h, scene._h, _stream_handle(stream)))

    def set_palette(self, palette: np.ndarray, stream=None) -> None:
        p = np.ascontiguousarray(palette, dtype=np.uint64)
        assert p.shape == (256,)
        L.check(L.lib().vrt_world_set_palette(self._h, p.ctypes.data, _stream_handle(stream)))

    def info(self) -> L.WorldInfo:
        i = L.WorldInfo()
        L.check(L.lib().vrt_world_get_info(self._h, C.byref(i)))
        return i

    def buffers(self) -> L.WorldBuffers:
        b = L.WorldBuffers()
        L.check(L.lib().vrt_world_get_buffers(self._h, C.byref(b)))
        return b

    def cast(self, rays: RayBatch, max_iters: int = 128, with_pos: bool = False, schedule: int = 0,
             chunk: int = 0, waves_per_cu: int = 0, counters: torch.Tensor | None = None,
             out: HitBatch | None = None, count: torch.Tensor | None = None, refill: int = 0,
             order: int = 0, with_uv: bool = False, with_material: bool = False, stream=None,
             pass_cap: int = 0) -> HitBatch:
        """Trace rays (RayCast, CpuRenderer.cpp:172-224).  count: optional device int32 holding the
        number of rays to trace (rays.n is then the capacity); order: the grid's block order
        (VRT_ORDER_FORWARD / VRT_ORDER_REVERSE, same results); pass_cap: the compact schedule's
        first-pass steps (schedule 3, 0 = 32)."""
        hits = out if out is not None else HitBatch.empty(rays.n, with_pos, rays.o.device, with_uv, with_material)
        r = rays.to_c()
        h = hits.to_c()
        opts = L.CastOpts(schedule, refill, chunk, waves_per_cu, counters.data_ptr() if counters is not None else 0,
                          count.data_ptr() if count is not None else 0, order, pass_cap)
        L.check(L.lib().vrt_cast_ex(self._h, C.byref(r), max_iters, C.byref(h), C.byref(opts),
                                    _stream_handle(stream)))
        return hits

    def cast_shader(self, rays: RayBatch, method: str = "full", ray_cell_lut: bool = False, max_iters: int = 0,
                    with_pos: bool = False, out: HitBatch | None = None, stream=None) -> HitBatch:
        """The GLSL traversal (VoxelTraversal.glsl rayCast / rayCastCoarse): method "full" or "coarse";
        ray_cell_lut masks with the ray-cell interaction table (u_UseAnisotropicLods); max_iters 0 is
        the shader's cap (256 / 96)."""
        m = {"full": L.VRT_SHADER_FULL, "coarse": L.VRT_SHADER_COARSE}[method]
        hits = out if out is not None else HitBatch.empty(rays.n, with_pos, rays.o.device)
        r = rays.to_c()
        h = hits.to_c()
        opts = L.ShaderOpts(m, int(ray_cell_lut), max_iters, 0)
        L.check(L.lib().vrt_cast_shader(self._h, C.byref(r), C.byref(opts), C.byref(h), _stream_handle(stream)))
        return hits

    def ray_cast(self, origin, direction, max_iters: int = 1024) -> HitResult:
        """HitResult VoxelMap::RayCast(dvec3 origin, dvec3 dir, uint32_t maxIters = 1024) (VoxelMap.h:214)."""
        o = (C.c_double * 3)(*origin)
        d = (C.c_double * 3)(*direction)
        h = L.HitF64()
        L.check(L.lib().vrt_cast_one_f64(self._h, C.byref(o), C.byref(d), max_iters, C.byref(h)))
        return HitResult(h.distance, tuple(h.normal), tuple(h.uv), tuple(h.voxel))

    def close(self) -> None:
        if self._h and self._h.value:
            L.check(L.lib().vrt_world_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cast_frame(world: "World", cam: "CameraSpec", max_iters: int = 128, seed: int = 1, with_pos: bool = False,
               schedule: int = 1, refill: int = 0, chunk: int = 0, counters: torch.Tensor | None = None,
               out=None, order: int = 0, stream=None):
    """vrt_cast_frame: (primary HitBatch, bounce HitBatch indexed by pixel); order: the grid's
    block order (VRT_ORDER_FORWARD / VRT_ORDER_REVERSE, same results)."""
    n = cam.n
    if out is None:
        dev = "cuda"
        out = (HitBatch.empty(n, with_pos, dev), HitBatch.empty(n, False, dev))
    hp, hb = out
    a, b = hp.to_c(), hb.to_c()
    o = L.FrameOpts(schedule, refill, chunk, counters.data_ptr() if counters is not None else 0, order)
    L.check(L.lib().vrt_cast_frame(world.handle, C.byref(cam.c), max_iters, seed, C.byref(a), C.byref(b),
                                   C.byref(o), _stream_handle(stream)))
    return hp, hb


class GBufferImages:
    """G-buffer images of vrt_render_frame on the device: albedo_normal [h, w] int32 (RGBA8 bytes),
    irradiance [h, w, 4] float32, depth [h, w] float32."""

    def __init__(self, width: int, height: int, device="cuda"):
        self.albedo_normal = torch.empty((height, width), dtype=torch.int32, device=device)
        self.irradiance = torch.empty((height, width, 4), dtype=torch.float32, device=device)
        self.depth = torch.empty((height, width), dtype=torch.float32, device=device)

    def to_c(self) -> L.GBuffer:
        return L.GBuffer(self.albedo_normal.data_ptr(), self.irradiance.data_ptr(), self.depth.data_ptr())


def render_frame(world: "World", cam: "CameraSpec", max_bounces: int = 1, frame: int = 0,
                 ray_cell_lut: bool = False, sky=(0.6, 0.75, 1.0), near_z: float = 0.0, far_z: float = 0.0,
                 out: GBufferImages | None = None, order: int = 0, stream=None) -> GBufferImages:
    """vrt_render_frame: one frame of the GPU renderer (VoxelRender.comp:28-94) into G-buffer images;
    order: the 8x8 tiles in row order or last tile first (VRT_ORDER_*, same images)."""
    g = out if out is not None else GBufferImages(cam.c.width, cam.c.height)
    o = L.RenderOpts(max_bounces, frame, int(ray_cell_lut), order, (C.c_float * 3)(*sky), near_z, far_z)
    c = g.to_c()
    L.check(L.lib().vrt_render_frame(world.handle, C.byref(cam.c), C.byref(o), C.byref(c), _stream_handle(stream)))
    return g


def release_stream(stream=None) -> None:
    """vrt_stream_release: frees the library's per-stream cast state (persistent queue counters,
    compact survivor state) of `stream` (torch's current stream when None), after synchronizing it."""
    L.check(L.lib().vrt_stream_release(_stream_handle(stream)))


def new_counters(device="cuda") -> torch.Tensor:
    """Zeroed vrt_cast_counters on the device (rays, iters, descents, hits)."""
    return torch.zeros(4, dtype=torch.int64, device=device)


def bounce(rays: RayBatch, hits: HitBatch, width: int, seed: int, stream=None, order: int = 0):
    """Diffuse bounce rays of the hits (vrt_bounce_generate, or vrt_bounce_generate_ex for order 1:
    grouped by direction octant); returns (RayBatch, src, count tensor)."""
    assert hits.pos is not None, "bounce generation needs hit positions"
    n = rays.n
    dev = rays.o.device
    out = RayBatch.empty(n, rays.world_origin, dev)
    src = torch.empty(n, dtype=torch.int32, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    scratch = torch.empty(int(L.lib().vrt_bounce_scratch_bytes(n)) // 4 + 1, dtype=torch.int32, device=dev)
    r = rays.to_c()
    h = hits.to_c()
    bufs = out.buffers()
    if order:
        L.check(L.lib().vrt_bounce_generate_ex(C.byref(r), C.byref(h), width, seed, order, C.byref(bufs),
                                               src.data_ptr(), count.data_ptr(), scratch.data_ptr(),
                                               _stream_handle(stream)))
    else:
        L.check(L.lib().vrt_bounce_generate(C.byref(r), C.byref(h), width, seed, C.byref(bufs), src.data_ptr(),
                                            count.data_ptr(), scratch.data_ptr(), _stream_handle(stream)))
    return out, src, count


# ------------------------------------------------------------------ batched


class Batch:
    """Camera segments (rays built in the kernel) followed by ray segments (read from `rays`),
    one hit range per segment in that order (vrt_batch).

    ray_segments: (ray_offset, count, world_origin, count_tensor_or_None); with a device count,
    `count` is the capacity.
    """

    def __init__(self, cameras: list, ray_segments: list = (), rays: RayBatch | None = None, max_iters: int = 128,
                 with_pos: bool = False, device="cuda"):
        self.cameras = list(cameras)
        self.ray_segments = list(ray_segments)
        self.rays = rays
        self.max_iters = max_iters
        segs = []
        off = 0
        self.ranges = []
        for k, cam in enumerate(self.cameras):
            segs.append(L.Segment(L.VRT_SEG_CAMERA, k, 0, off, cam.n, (C.c_int32 * 3)(0, 0, 0), 0))
            self.ranges.append((off, cam.n))
            off += cam.n
        for (ro, cnt, W, cptr) in self.ray_segments:
            segs.append(L.Segment(L.VRT_SEG_RAYS, 0, ro, off, cnt, (C.c_int32 * 3)(*W),
                                  cptr.data_ptr() if cptr is not None else 0))
            self.ranges.append((off, cnt))
            off += cnt
        self.total = off
        self.hits = HitBatch.empty(off, with_pos, device)
        self._cams = (L.Camera * max(1, len(self.cameras)))(*[c.c for c in self.cameras])
        self._segs = (L.Segment * max(1, len(segs)))(*segs)
        self._nseg = len(segs)

    def to_c(self) -> L.Batch:
        r = self.rays.to_c() if self.rays is not None else L.Rays()
        return L.Batch(C.cast(self._cams, C.c_void_p), len(self.cameras), C.cast(self._segs, C.c_void_p),
                       self._nseg, r, self.hits.to_c(), self.max_iters)

    def segment_hits(self, k: int):
        off, cnt = self.ranges[k]
        return (self.hits.t[off:off + cnt], self.hits.rec[off:off + cnt],
                self.hits.pos[:, off:off + cnt] if self.hits.pos is not None else None)


class BatchPlan:
    """vrt_batch_plan: device tables for one batch and tuning; launch or graph-launch it."""

    def __init__(self, world: World, batch: Batch, tuning=(8, 8, 1, 0), stream=None):
        self.batch = batch
        self._b = batch.to_c()
        t = L.BatchTuning(*tuning)
        out = C.c_void_p()
        L.check(L.lib().vrt_batch_plan_create(world.handle, C.byref(self._b), C.byref(t), _stream_handle(stream),
                                              C.byref(out)))
        self._h = out

    def launch(self, counters: torch.Tensor | None = None, stream=None) -> None:
        L.check(L.lib().vrt_batch_launch(self._h, counters.data_ptr() if counters is not None else 0,
                                         _stream_handle(stream)))

    def graph_launch(self, stream=None) -> None:
        L.check(L.lib().vrt_batch_graph_launch(self._h, _stream_handle(stream)))

    def close(self) -> None:
        if self._h and self._h.value:
            L.check(L.lib().vrt_batch_plan_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(world: World, batch: Batch, cache_path: str | None = None, stream=None):
    """vrt_batch_autotune: returns ((tile_w, tile_h, rays_per_lane, chunk), ms)."""
    b = batch.to_c()
    best = L.BatchTuning()
    ms = C.c_float()
    L.check(L.lib().vrt_batch_autotune(world.handle, C.byref(b), cache_path.encode() if cache_path else None,
                                       C.byref(best), C.byref(ms), _stream_handle(stream)))
    return (best.tile_w, best.tile_h, best.rays_per_lane, best.chunk), ms.value


# ------------------------------------------------------------------ differentiable

SUN_DIR = (0.3, 0.9, -0.28)   # VoxelRender.comp:133, normalised below
SUN_INTENSITY = 5.0           # VoxelRender.comp:134


def sun_direction() -> tuple:
    x, y, z = SUN_DIR
    n = (x * x + y * y + z * z) ** 0.5
    return (x / n, y / n, z / n)


_torch_ext = None


def torch_ext():
    """The libtorch autograd extension (_vrt_torch.so, built next to libvoxelrt_hip.so)."""
    global _torch_ext
    if _torch_ext is None:
        import importlib.util
        from pathlib import Path

        path = Path(__file__).resolve().parent / "_vrt_torch.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it with `python -m voxelrt_amd.build`")
        L.lib()  # load libvoxelrt_hip.so first
        spec = importlib.util.spec_from_file_location("_vrt_torch", str(path))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _torch_ext = mod
    return _torch_ext


def _pack_cameras(cameras) -> torch.Tensor:
    raw = b"".join(bytes(c.c) for c in cameras)
    if not raw:
        return torch.empty(0, dtype=torch.uint8)
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8)


GRAD_MODES = {"bricks": 0, "atomic": 1, "sparse": 2}


def render(world: World, albedo: torch.Tensor, cameras, max_iters: int = 128, light=None,
           intensity: float = SUN_INTENSITY, stream=None, grad: str = "bricks") -> torch.Tensor:
    """Differentiable Lambertian render: [pixels, 3] image, gradients flow to `albedo`
    ([voxel slots * 512, 3] fp32) through the libtorch autograd Function.  grad: the albedo
    gradient's form, "bricks" (dense, written whole: vrt_render_bwd_bricks), "atomic" (dense,
    zero-filled and scatter-added: vrt_render_bwd) or "sparse" (a COO tensor with one unsummed
    entry per pixel: vrt_render_bwd_sparse)."""
    light = light if light is not None else sun_direction()
    return torch_ext().render(albedo, world.handle.value, _pack_cameras(cameras), max_iters, list(light),
                              float(intensity), _stream_handle(stream), GRAD_MODES[grad])


def render_saved(world: World, albedo: torch.Tensor, cameras, max_iters: int = 128, light=None,
                 intensity: float = SUN_INTENSITY, stream=None):
    """Forward render returning (image, saved voxel ids int32 (-1 = miss), saved weights)."""
    light = light if light is not None else sun_direction()
    return torch_ext().render_saved(albedo, world.handle.value, _pack_cameras(cameras), max_iters, list(light),
                                    float(intensity), _stream_handle(stream))


def render_backward(world: World, grad_image: torch.Tensor, saved_voxel: torch.Tensor, saved_weight: torch.Tensor,
                    nvox: int, stream=None, mode: str = "atomic"):
    """The albedo gradient [nvox, 3] from saved hits: mode "atomic" (vrt_render_bwd into a fresh
    zero gradient), "bricks" (vrt_render_bwd_bricks, written whole) or "sparse"
    (vrt_render_bwd_sparse: returns (index int64 [n], values [n, 3]), one entry per pixel)."""
    need = world.info().slot_high_water * 512
    if nvox < need:
        raise ValueError(f"render_backward: nvox {nvox} < slot_high_water * 512 = {need}")
    if saved_voxel.numel() and grad_image.numel() != 3 * saved_voxel.numel():
        raise ValueError("render_backward: grad_image must hold 3 values per saved pixel")
    g = grad_image.contiguous()
    lib, n, sh = L.lib(), saved_voxel.numel(), _stream_handle(stream)
    if mode == "sparse":
        index = torch.empty(n, dtype=torch.int64, device=grad_image.device)
        values = torch.empty((n, 3), dtype=torch.float32, device=grad_image.device)
        L.check(lib.vrt_render_bwd_sparse(world.handle, n, g.data_ptr(), saved_voxel.data_ptr(),
                                          saved_weight.data_ptr(), index.data_ptr(), values.data_ptr(), sh))
        return index, values
    if mode == "bricks":
        grad = torch.empty((nvox, 3), dtype=torch.float32, device=grad_image.device)
        nbytes = int(lib.vrt_render_bwd_workspace_bytes(n, nvox))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=grad_image.device)
        L.check(lib.vrt_render_bwd_bricks(world.handle, n, g.data_ptr(), saved_voxel.data_ptr(), saved_weight.data_ptr(),
                                          nvox, grad.data_ptr(), ws.data_ptr(), nbytes, sh))
        return grad
    grad = torch.zeros((nvox, 3), dtype=torch.float32, device=grad_image.device)
    L.check(lib.vrt_render_bwd(world.handle, n, g.data_ptr(), saved_voxel.data_ptr(), saved_weight.data_ptr(),
                               grad.data_ptr(), sh))
    return grad
