"""GPU parity of the shader method (vrt_cast_shader) against the oracle's restatement of
VoxelTraversal.glsl, bit-exact, for rayCast / rayCastCoarse with and without the ray-cell table."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import assert_hits_equal, blob_world, random_rays
from test_oracle import dense_world
from test_shader_oracle import camera_rays

pytestmark = pytest.mark.gpu

MODES = [("full", False), ("full", True), ("coarse", False), ("coarse", True)]


def gpu_shader(world, o, d, W, method, lut, max_iters=0):
    from voxelrt_amd import RayBatch

    h = world.cast_shader(RayBatch.from_numpy(o, d, W), method, lut, max_iters, with_pos=True)
    torch.cuda.synchronize()
    return h.t.cpu().numpy(), h.rec.cpu().numpy(), h.pos.cpu().numpy()


def check_mode(world, ow, o, d, W, method, lut, what, max_iters=0):
    t, rec, pos = gpu_shader(world, o, d, W, method, lut, max_iters)
    ref = ow.cast_shader(o, d, W, method, lut, max_iters, with_pos=True)
    assert_hits_equal(t, rec, ref["t"], ref["rec"], f"{what} {method} lut={lut}")
    assert np.array_equal(t.view(np.uint32), ref["t"].view(np.uint32))
    assert np.array_equal(pos.view(np.uint32), ref["pos"].view(np.uint32))
    return rec


def test_c1_terrain_camera(vrt, oracle_mod, cuda):
    from voxelrt_amd import workloads as W
    import oracle

    sc = W.C1_WORLD.build()
    world = vrt.World.from_scene(sc)
    ow = oracle.OracleWorld.from_stream(sc.stream())
    cam = W.c1_camera()
    o, d = cam.rays_host()
    parity = ow.cast(o, d, cam.world_origin, 256)
    for method, lut in MODES:
        rec = check_mode(world, ow, o, d, cam.world_origin, method, lut, "C1")
        hit = (rec[:, 3] & 0xFF) != 0
        same = np.all(rec[:, :3] == parity["rec"][:, :3], axis=1) & (hit == ((parity["rec"][:, 3] & 0xFF) != 0))
        it = (rec[:, 3] >> 16) & 0xFFFF
        print(f"C1 {method} lut={lut}: hit voxel agreement with vrt_cast {same.mean():.4f}, "
              f"mean iterations {it.mean():.2f}", flush=True)
        if method == "full":
            assert same.mean() > 0.99


@pytest.mark.parametrize("seed,size", [(1, 64), (2, 128), (3, 256)])
def test_random_world_cameras(vrt, oracle_mod, cuda, seed, size):
    vox = blob_world(size, seed=seed, density=0.001, blobs=6)
    if size >= 256:
        vox[:128, :128, :128] = 0
    world = vrt.World.from_scene(vrt.Scene.from_dense(vox))
    ow = dense_world(oracle_mod, vox)
    rng = np.random.default_rng(seed + 100)
    for k in range(3):
        o, d, W = camera_rays(rng, vox, 4000)
        for method, lut in MODES:
            check_mode(world, ow, o, d, W, method, lut, f"seed={seed} cam={k}")


def test_clipped_and_special_rays(vrt, oracle_mod, cuda):
    """Origins outside the box (clipRayToAABB), large W, short caps, denormal and huge directions."""
    vox = blob_world(128, seed=7, density=0.003, blobs=5)
    world = vrt.World.from_scene(vrt.Scene.from_dense(vox))
    ow = dense_world(oracle_mod, vox)
    o, d = random_rays(20000, 128, seed=8, inside=False)
    o = o * np.float32(1.5) - np.float32(32.0)
    d[0, :200] = np.float32(1e-40)
    d[1, 200:400] = np.float32(1e30)
    d[2, 400:600] = np.float32(-0.0) + np.float32(2.0**-24)
    for W in [(0, 0, 0), (40, -3, 17)]:
        oo = (o - np.array(W, np.float32)[:, None]).astype(np.float32)
        for method, lut in MODES:
            for cap in (0, 7):
                check_mode(world, ow, oo, d, W, method, lut, f"clip W={W} cap={cap}", cap)


@pytest.mark.parametrize("size,ysize,seed", [(128, 64, 41), (64, 128, 42), (256, 32, 43)])
def test_non_cubic_world_shader_parity(vrt, oracle_mod, cuda, size, ysize, seed):
    """Worlds whose y extent differs from x/z: the two-compare bounds test, clipping to a
    non-cubic box and sector-of-sectors groups that the world's y extent cuts short
    (a 32-voxel-high world has one sector layer in groups four sectors high)."""
    from helpers import random_rays as rr

    vox = blob_world(size, seed=seed, density=0.004, blobs=5, ysize=ysize)
    world = vrt.World.from_scene(vrt.Scene.from_dense(vox))
    ow = dense_world(oracle_mod, vox)
    info = world.info()
    assert (info.size_xz, info.size_y) == (size, ysize)
    for inside in (True, False):
        o, d = rr(8000, size, seed=seed + 100, ysize=ysize, inside=inside)
        for W in [(0, 0, 0), (32, -64, 96)]:
            oo = (o - np.array(W, np.float32)[:, None]).astype(np.float32)
            for method, lut in MODES:
                rec = check_mode(world, ow, oo, d, W, method, lut, f"{size}x{ysize} inside={inside} W={W}")
                assert ((rec[:, 3] & 0xFF) != 0).mean() > 0.05


@pytest.fixture(scope="module")
def c2_pair(vrt, oracle_mod):
    from voxelrt_amd import workloads as W

    sc = W.C2_WORLD.build()
    return vrt.World.from_scene(sc), oracle_mod.OracleWorld.from_stream(sc.stream())


@pytest.mark.parametrize("anchor", ["sector", "floor"])
def test_c2_camera_full_size(vrt, cuda, c2_pair, anchor):
    """The 1024^3 terrain + caves world and its 1920x1080 camera: every ray of all four modes
    bit-exact with the oracle (records, t, hit positions), at both ray anchors."""
    from voxelrt_amd import workloads as W

    world, ow = c2_pair
    cam = W.c2_camera(anchor=anchor)
    o, d = cam.rays_host()
    for method, lut in MODES:
        rec = check_mode(world, ow, o, d, cam.world_origin, method, lut, f"C2 {anchor}")
        assert ((rec[:, 3] & 0xFF) != 0).mean() > 0.3
