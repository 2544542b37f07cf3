"""Host-side checks of libvoxelrt_hip.so that need no GPU: the C ABI exports,
error handling without a device, the seeded generators and host ray builders."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions() -> list[str]:
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w]+\s*\**\s+\**(vrt_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(vrt):
    names = declared_functions()
    assert len(names) >= 30
    lib = vrt.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in include/*.h but not exported: {missing}"


def test_abi_version_and_error_string(vrt):
    lib = vrt.lib()
    assert lib.vrt_abi_version() == 2
    assert isinstance(lib.vrt_last_error(), bytes)


def test_null_arguments_are_rejected(vrt):
    from voxelrt_amd import _lib as L

    lib = vrt.lib()
    assert lib.vrt_world_create(None, None, None) == L.VRT_ERR_INVALID_ARGUMENT
    assert b"null" in lib.vrt_last_error()
    assert lib.vrt_world_destroy(None) == L.VRT_ERR_INVALID_ARGUMENT
    assert lib.vrt_cast(None, None, 128, None, None) == L.VRT_ERR_INVALID_ARGUMENT
    assert lib.vrt_world_sync(None, 0, None, None, None, None) == L.VRT_ERR_INVALID_ARGUMENT
    desc = L.WorldDesc(4, 8, 0, 0)
    out = C.c_void_p()
    assert lib.vrt_world_create(C.byref(desc), None, C.byref(out)) == L.VRT_ERR_INVALID_ARGUMENT
    assert b"5..14" in lib.vrt_last_error()
    assert lib.vrt_world_sync_bricks(None, 0, None, None, None, None, None) == L.VRT_ERR_INVALID_ARGUMENT


def test_bounce_generation_rejects_a_wrapping_primary_count(vrt):
    """n above 2^32 - 4096 would wrap the 32-bit block counts: rejected before any launch (the
    pointers are placeholders that are never dereferenced)."""
    from voxelrt_amd import _lib as L

    lib = vrt.lib()
    fake = 0x1000
    prim = L.Rays(fake, fake, fake, fake, fake, fake, (L.i32 * 3)(0, 0, 0), 0xFFFFFFFF)
    hits = L.Hits(fake, fake, fake)
    out = L.RayBuffers(fake, fake, fake, fake, fake, fake)
    rc = lib.vrt_bounce_generate(C.byref(prim), C.byref(hits), 1920, 1, C.byref(out), fake, fake, fake, None)
    assert rc == L.VRT_ERR_INVALID_ARGUMENT
    assert b"2^32 - 4096" in lib.vrt_last_error()
    # the scratch size is computed in 64 bits: one u32 per 256 primaries, plus one
    assert lib.vrt_bounce_scratch_bytes(0xFFFFFFFF) == ((0xFFFFFFFF + 255) // 256 + 1) * 4


def test_terrain_generator_is_deterministic(vrt):
    from voxelrt_amd import workloads as W

    a = W.C1_WORLD.build(threads=1).stream()
    b = W.C1_WORLD.build(threads=4).stream()
    assert a["nbricks"] == b["nbricks"] > 0
    assert np.array_equal(a["sector_idx"], b["sector_idx"])
    assert np.array_equal(a["alloc_mask"], b["alloc_mask"])
    assert np.array_equal(a["bricks"], b["bricks"])
    assert np.all(np.diff(a["sector_idx"].astype(np.int64)) > 0)
    counts = [bin(int(m)).count("1") for m in a["alloc_mask"]]
    assert sum(counts) == a["nbricks"]


def test_scene_from_dense_roundtrip(vrt):
    from helpers import blob_world

    vox = blob_world(64, seed=3)
    sc = vrt.Scene.from_dense(vox)
    rng = np.random.default_rng(0)
    for p in rng.integers(0, 64, size=(4000, 3)):
        assert sc.voxel(*p) == vox[p[1], p[2], p[0]]
    s = sc.stream()
    assert s["solid_voxels"] == int((vox != 0).sum())
    # every listed brick holds at least one solid voxel
    br = s["bricks"].reshape(-1, 512)
    assert np.all(br.any(axis=1))


def test_spheres_generator_fill(vrt):
    sc = vrt.Scene.spheres(7, 7, 20, 4.0, 12.0, seed=9)
    s = sc.stream()
    assert 0 < s["solid_voxels"] < 128**3


def test_camera_rays_host_shape_and_tiling(vrt):
    cam = vrt.CameraSpec((32.5, 40.25, 8.75), 10.0, -20.0, 64, 32, tile=(8, 8))
    o, d = cam.rays_host()
    assert o.shape == (3, 64 * 32)
    assert np.allclose(np.linalg.norm(d.astype(np.float64), axis=0), 1.0, atol=1e-6)
    assert np.all(d != 0)
    assert cam.world_origin == (32, 40, 8)
    assert np.all(o[0] == np.float32(0.5)) and np.all(o[1] == np.float32(0.25))
    # tiled order is a permutation of the row-major order
    cam_rm = vrt.CameraSpec((32.5, 40.25, 8.75), 10.0, -20.0, 64, 32, tile=(1, 1))
    _, d_rm = cam_rm.rays_host()
    key = lambda a: np.lexsort(a)  # noqa: E731
    assert np.array_equal(np.sort(d.view(np.uint32), axis=1)[:, key(np.sort(d.view(np.uint32), axis=1))],
                          np.sort(d_rm.view(np.uint32), axis=1)[:, key(np.sort(d_rm.view(np.uint32), axis=1))])
    # first 64 rays of the 8x8 tiling are pixels (0..7, 0..7)
    pix = [(i % 8, i // 8) for i in range(64)]
    rm_index = [py * 64 + px for px, py in pix]
    assert np.array_equal(d[:, :64], d_rm[:, rm_index])
    with pytest.raises(vrt.VrtError):
        vrt.CameraSpec((0, 0, 0), 0, 0, 60, 32, tile=(8, 8))


def test_world_create_without_device_reports_no_device(vrt):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(vrt.VrtError) as e:
        vrt.World(8, 8)
    assert e.value.status == 4


def test_oversized_launches_are_rejected_before_any_device_work(vrt):
    """Counts whose 32-bit grid (blocks x lanes) would wrap are refused by the
    argument checks, which run before the world is read or a kernel launched."""
    import torch

    from voxelrt_amd import _lib as L

    if torch.cuda.is_available():
        pytest.skip("a device is present: the pointers below are host placeholders")
    lib = vrt.lib()
    # a zeroed host block larger than a vrt_world: the checks read only its failed-state flag
    # (false) before rejecting the sizes; no device memory is touched
    host = C.create_string_buffer(1 << 16)
    p = C.addressof(host)
    fake_world = C.c_void_p(p)
    # render backward: at most 2^32 - 256 pixels
    assert lib.vrt_render_bwd(fake_world, 0xFFFFFFFF, p, p, p, p, None) == L.VRT_ERR_INVALID_ARGUMENT
    assert b"2^32 - 256" in lib.vrt_last_error()
    # camera rays: 65536 x 65536 pixels is 2^32
    cam = L.Camera()
    cam.width = cam.height = 65536
    cam.tile_w = cam.tile_h = 1
    bufs = L.RayBuffers(p, p, p, p, p, p)
    assert lib.vrt_camera_rays(C.byref(cam), C.byref(bufs), None) == L.VRT_ERR_INVALID_ARGUMENT
    assert b"too many pixels" in lib.vrt_last_error()
    # shader cast: at most 2^32 - 64 rays
    rays = L.Rays(p, p, p, p, p, p)
    rays.n = 0xFFFFFFF0
    hits = L.Hits(p, (p + 15) & ~15, 0)
    opts = L.ShaderOpts(0, 0, 0, 0)
    assert lib.vrt_cast_shader(fake_world, C.byref(rays), C.byref(opts), C.byref(hits), None) == \
        L.VRT_ERR_INVALID_ARGUMENT
    assert b"too many rays" in lib.vrt_last_error()
    # FP64 picking cast: its loop runs every step up to the cap, so the cap is 1..65535
    o3, d3 = (C.c_double * 3)(0.5, 0.5, 0.5), (C.c_double * 3)(1.0, 0.0, 0.0)
    hit = L.HitF64()
    for cap in (0, 65536, 0xFFFFFFFF):
        assert lib.vrt_cast_one_f64(fake_world, C.byref(o3), C.byref(d3), cap, C.byref(hit)) == \
            L.VRT_ERR_INVALID_ARGUMENT
        assert b"1..65535" in lib.vrt_last_error()

