"""ctypes wrapper of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  Never used by the product.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"

P = C.c_void_p
u32 = C.c_uint32


class OracleHitF64(C.Structure):
    _fields_ = [("distance", C.c_double), ("normal", C.c_float * 3), ("uv", C.c_float * 2),
                ("voxel", C.c_int32 * 3)]


class OracleCounters(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("iters", C.c_uint64), ("descents", C.c_uint64), ("hits", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = C.CDLL(str(LIB_PATH))
        h.oracle_world_create.argtypes = [u32, u32, u32, P, P, P, C.POINTER(P)]
        h.oracle_world_create.restype = C.c_int
        h.oracle_world_destroy.argtypes = [P]
        h.oracle_world_set_palette.argtypes = [P, P]
        h.oracle_world_get_voxel.argtypes = [P, C.c_int32, C.c_int32, C.c_int32]
        h.oracle_world_get_voxel.restype = C.c_uint8
        h.oracle_cast.argtypes = [P, C.POINTER(P * 3), C.POINTER(P * 3), C.POINTER(C.c_int32 * 3), u32, u32, u32, P,
                                  P, P, P, C.POINTER(P * 2)]
        h.oracle_count.argtypes = [P, C.POINTER(P * 3), C.POINTER(P * 3), C.POINTER(C.c_int32 * 3), u32, u32,
                                   C.POINTER(OracleCounters)]
        h.oracle_cast_simd.argtypes = [P, C.POINTER(P * 3), C.POINTER(P * 3), C.POINTER(C.c_int32 * 3), u32, u32, u32,
                                       P, P, C.POINTER(P * 3)]
        h.oracle_step_pos.argtypes = [P, u32, P, P, P, P, P]
        h.oracle_cast_simd.restype = C.c_int
        h.oracle_simd_width.restype = C.c_int
        h.oracle_num_cores.restype = C.c_int
        h.oracle_cast_f64.argtypes = [P, C.POINTER(C.c_double * 3), C.POINTER(C.c_double * 3), u32,
                                      C.POINTER(OracleHitF64)]
        h.oracle_dda_f64.argtypes = [P, C.POINTER(C.c_double * 3), C.POINTER(C.c_double * 3),
                                     C.POINTER(C.c_int32 * 3), u32, C.POINTER(C.c_int32 * 3),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
        h.oracle_dda_f64.restype = C.c_int
        h.oracle_pcg_init.argtypes = [u32, u32, u32]
        h.oracle_pcg_init.restype = u32
        h.oracle_pcg_next.argtypes = [C.POINTER(u32)]
        h.oracle_pcg_next.restype = u32
        h.oracle_sincos_turns.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        h.oracle_bounce.argtypes = [u32, P, P, C.POINTER(P * 3), C.POINTER(P * 3), u32, u32, C.POINTER(P * 3),
                                    C.POINTER(P * 3), P]
        h.oracle_bounce.restype = u32
        h.oracle_render_backward.argtypes = [u32, P, P, P, P, P]
        h.oracle_render_loss.argtypes = [u32, P, P, P, P]
        h.oracle_render_loss.restype = C.c_double
        h.oracle_ray_cell_lut.argtypes = [P]
        h.oracle_render_frame.argtypes = [P, u32, u32, C.POINTER(P * 3), C.POINTER(P * 3), C.POINTER(C.c_int32 * 3),
                                          C.POINTER(C.c_float * 3), C.POINTER(C.c_float * 3), C.c_float, C.c_float,
                                          u32, u32, u32, C.POINTER(C.c_float * 3), P, P, P]
        h.oracle_cast_shader.argtypes = [P, C.POINTER(P * 3), C.POINTER(P * 3), C.POINTER(C.c_int32 * 3), u32, u32,
                                         u32, u32, u32, P, P, C.POINTER(P * 3)]
        _lib = h
    return _lib


def _p3(a: np.ndarray):
    return (P * 3)(*[a[i].ctypes.data for i in range(3)])


class OracleWorld:
    """The reference's FlatVoxelStorage restated (dense per sector)."""

    def __init__(self, log2_xz: int, log2_y: int, sector_idx, alloc_mask, bricks):
        si = np.ascontiguousarray(sector_idx, dtype=np.uint32)
        am = np.ascontiguousarray(alloc_mask, dtype=np.uint64)
        br = np.ascontiguousarray(bricks, dtype=np.uint8)
        out = P()
        rc = lib().oracle_world_create(log2_xz, log2_y, len(si), si.ctypes.data, am.ctypes.data,
                                       br.ctypes.data if br.size else None, C.byref(out))
        if rc != 0:
            raise RuntimeError(f"oracle_world_create failed ({rc})")
        self._h = out
        self.size_xz = 1 << log2_xz
        self.size_y = 1 << log2_y

    @staticmethod
    def from_stream(s: dict) -> "OracleWorld":
        return OracleWorld(s["log2_size_xz"], s["log2_size_y"], s["sector_idx"], s["alloc_mask"], s["bricks"])

    def voxel(self, x, y, z) -> int:
        return int(lib().oracle_world_get_voxel(self._h, x, y, z))

    def cast(self, o: np.ndarray, d: np.ndarray, world_origin=(0, 0, 0), max_iters=128, begin=0, end=None,
             with_pos=False, with_descents=False, with_uv=False):
        """Scalar cast; returns dict(t, rec[n,4] int32, pos[3,n] | None, descents | None, uv[2,n] | None)
        for rays [begin, end)."""
        o = np.ascontiguousarray(o, dtype=np.float32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        n = o.shape[1]
        end = n if end is None else end
        t = np.zeros(n, dtype=np.float32)
        rec = np.zeros((n, 4), dtype=np.int32)
        pos = np.zeros((3, n), dtype=np.float32) if with_pos else None
        desc = np.zeros(n, dtype=np.uint16) if with_descents else None
        uv = np.zeros((2, n), dtype=np.float32) if with_uv else None
        W = (C.c_int32 * 3)(*world_origin)
        uvp = (P * 2)(uv[0].ctypes.data, uv[1].ctypes.data) if with_uv else None
        lib().oracle_cast(self._h, C.byref(_p3(o)), C.byref(_p3(d)), C.byref(W), begin, end, max_iters,
                          t.ctypes.data, rec.ctypes.data, C.byref(_p3(pos)) if with_pos else None,
                          desc.ctypes.data if with_descents else None, C.byref(uvp) if with_uv else None)
        return {"t": t[begin:end], "rec": rec[begin:end], "pos": pos[:, begin:end] if with_pos else None,
                "descents": desc[begin:end] if with_descents else None, "uv": uv[:, begin:end] if with_uv else None}

    def count(self, o, d, world_origin=(0, 0, 0), max_iters=128) -> dict:
        o = np.ascontiguousarray(o, dtype=np.float32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        c = OracleCounters()
        W = (C.c_int32 * 3)(*world_origin)
        lib().oracle_count(self._h, C.byref(_p3(o)), C.byref(_p3(d)), C.byref(W), o.shape[1], max_iters, C.byref(c))
        return {"rays": c.rays, "iters": c.iters, "descents": c.descents, "hits": c.hits}

    def cast_simd(self, o, d, world_origin=(0, 0, 0), max_iters=128, threads=0, with_pos=False):
        o = np.ascontiguousarray(o, dtype=np.float32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        n = o.shape[1]
        t = np.zeros(n, dtype=np.float32)
        rec = np.zeros((n, 4), dtype=np.int32)
        pos = np.zeros((3, n), dtype=np.float32) if with_pos else None
        W = (C.c_int32 * 3)(*world_origin)
        lanes = lib().oracle_cast_simd(self._h, C.byref(_p3(o)), C.byref(_p3(d)), C.byref(W), n, max_iters, threads,
                                       t.ctypes.data, rec.ctypes.data, C.byref(_p3(pos)) if with_pos else None)
        if lanes == 0:
            raise RuntimeError("host has neither AVX-512 nor AVX2+FMA")
        return {"t": t, "rec": rec, "pos": pos, "lanes": lanes}

    def step_pos(self, v: np.ndarray, d: np.ndarray):
        """GetStepPos alone for in-bounds voxels v [n, 3] and directions d [n, 3]: (v_out, solid, descended)."""
        v = np.ascontiguousarray(v, dtype=np.int32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        n = v.shape[0]
        out = np.zeros((n, 3), np.int32)
        solid = np.zeros(n, np.uint8)
        desc = np.zeros(n, np.uint8)
        lib().oracle_step_pos(self._h, n, v.ctypes.data, d.ctypes.data, out.ctypes.data, solid.ctypes.data,
                              desc.ctypes.data)
        return out, solid.astype(bool), desc.astype(bool)

    def cast_shader(self, o, d, world_origin=(0, 0, 0), method="full", lut=False, max_iters=0, begin=0, end=None,
                    with_pos=False) -> dict:
        """Shader-method cast (VoxelTraversal.glsl rayCast / rayCastCoarse) of rays [begin, end)."""
        o = np.ascontiguousarray(o, dtype=np.float32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        n = o.shape[1]
        end = n if end is None else end
        t = np.zeros(n, dtype=np.float32)
        rec = np.zeros((n, 4), dtype=np.int32)
        pos = np.zeros((3, n), dtype=np.float32) if with_pos else None
        W = (C.c_int32 * 3)(*world_origin)
        m = {"full": 0, "coarse": 1}[method]
        lib().oracle_cast_shader(self._h, C.byref(_p3(o)), C.byref(_p3(d)), C.byref(W), begin, end, m, int(lut),
                                 max_iters, t.ctypes.data, rec.ctypes.data, C.byref(_p3(pos)) if with_pos else None)
        return {"t": t[begin:end], "rec": rec[begin:end], "pos": pos[:, begin:end] if with_pos else None}

    def render_frame(self, width, height, o, d, world_origin, forward, sun, pa, pb, max_bounces, frame, lut, sky):
        """G-buffer of one frame (VoxelRender.comp:28-94 over the shader method) for row-major primary rays."""
        o = np.ascontiguousarray(o, dtype=np.float32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        an = np.zeros((height, width), dtype=np.uint32)
        irr = np.zeros((height, width, 4), dtype=np.float32)
        dep = np.zeros((height, width), dtype=np.float32)
        f3 = C.c_float * 3
        lib().oracle_render_frame(self._h, width, height, C.byref(_p3(o)), C.byref(_p3(d)),
                                  C.byref((C.c_int32 * 3)(*world_origin)), C.byref(f3(*forward)), C.byref(f3(*sun)),
                                  pa, pb, max_bounces, frame, int(lut), C.byref(f3(*sky)), an.ctypes.data,
                                  irr.ctypes.data, dep.ctypes.data)
        return {"albedo_normal": an, "irradiance": irr, "depth": dep}

    def set_palette(self, palette) -> None:
        pal = np.ascontiguousarray(palette, dtype=np.uint64)
        lib().oracle_world_set_palette(self._h, pal.ctypes.data)

    def cast_f64(self, origin, direction, max_iters=1024) -> dict:
        o = (C.c_double * 3)(*origin)
        d = (C.c_double * 3)(*direction)
        h = OracleHitF64()
        lib().oracle_cast_f64(self._h, C.byref(o), C.byref(d), max_iters, C.byref(h))
        return {"distance": h.distance, "normal": tuple(h.normal), "uv": tuple(h.uv), "voxel": tuple(h.voxel)}

    def dda(self, o, d, world_origin=(0, 0, 0), max_steps=4096):
        oo = (C.c_double * 3)(*o)
        dd = (C.c_double * 3)(*d)
        W = (C.c_int32 * 3)(*world_origin)
        v = (C.c_int32 * 3)()
        t = C.c_double()
        gap = C.c_double()
        hit = lib().oracle_dda_f64(self._h, C.byref(oo), C.byref(dd), C.byref(W), max_steps, C.byref(v), C.byref(t),
                                   C.byref(gap))
        return bool(hit), tuple(v), t.value, gap.value

    def close(self):
        if self._h and self._h.value:
            lib().oracle_world_destroy(self._h)
            self._h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ray_cell_lut() -> np.ndarray:
    """The ray-cell interaction LUT (GpuRenderer.cpp:192-210), [8 octants * 64 cells] u64."""
    t = np.zeros(512, dtype=np.uint64)
    lib().oracle_ray_cell_lut(t.ctypes.data)
    return t


def simd_width() -> int:
    return int(lib().oracle_simd_width())


def num_cores() -> int:
    return int(lib().oracle_num_cores())


def pcg_init(px, py, seed) -> int:
    return int(lib().oracle_pcg_init(px, py, seed))


def pcg_next(state: int) -> tuple[int, int]:
    s = u32(state)
    v = lib().oracle_pcg_next(C.byref(s))
    return int(v), int(s.value)


def sincos_turns(u: float) -> tuple[float, float]:
    s = C.c_float()
    c = C.c_float()
    lib().oracle_sincos_turns(u, C.byref(s), C.byref(c))
    return s.value, c.value


def bounce(rec: np.ndarray, pos: np.ndarray, d: np.ndarray, width: int, seed: int):
    n = rec.shape[0]
    rec = np.ascontiguousarray(rec, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    d = np.ascontiguousarray(d, dtype=np.float32)
    t = np.zeros(n, dtype=np.float32)
    oo = np.zeros((3, n), dtype=np.float32)
    od = np.zeros((3, n), dtype=np.float32)
    src = np.zeros(n, dtype=np.uint32)
    k = lib().oracle_bounce(n, t.ctypes.data, rec.ctypes.data, C.byref(_p3(pos)), C.byref(_p3(d)), width, seed,
                            C.byref(_p3(oo)), C.byref(_p3(od)), src.ctypes.data)
    return oo[:, :k].copy(), od[:, :k].copy(), src[:k].copy()


def render_backward(saved_voxel, saved_weight, grad_image, nvox):
    """FP64 backward: (grad [nvox, 3], abssum [nvox, 3])."""
    sv = np.ascontiguousarray(saved_voxel, dtype=np.uint32)
    sw = np.ascontiguousarray(saved_weight, dtype=np.float32)
    gi = np.ascontiguousarray(grad_image, dtype=np.float32)
    g = np.zeros(nvox * 3, dtype=np.float64)
    a = np.zeros(nvox * 3, dtype=np.float64)
    lib().oracle_render_backward(len(sv), sv.ctypes.data, sw.ctypes.data, gi.ctypes.data, g.ctypes.data,
                                 a.ctypes.data)
    return g.reshape(-1, 3), a.reshape(-1, 3)


def render_loss(saved_voxel, saved_weight, albedo64, target) -> float:
    sv = np.ascontiguousarray(saved_voxel, dtype=np.uint32)
    sw = np.ascontiguousarray(saved_weight, dtype=np.float32)
    al = np.ascontiguousarray(albedo64, dtype=np.float64)
    tg = np.ascontiguousarray(target, dtype=np.float32)
    return float(lib().oracle_render_loss(len(sv), sv.ctypes.data, sw.ctypes.data, al.ctypes.data, tg.ctypes.data))
