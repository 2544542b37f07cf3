"""GPU parity of the plain per-ray cast (vrt_cast through the C ABI) against the oracle."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

from helpers import assert_hits_equal, blob_world, random_rays
from test_oracle import dense_world

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
# (schedule, parameter): grid in forward and reverse block order; persistent with 64/256/1024-ray
# chunks; compact with first passes of 1, 7 and 32 steps (1: every survivor resumes after its
# first step)
SCHEDULES = [(1, 0), (-1, 0), (2, 64), (2, 256), (2, 1024), (3, 1), (3, 7), (3, 32)]


def gpu_cast(world, o, d, W, max_iters=128, schedule=1, chunk=0, with_pos=False, counters=None):
    from voxelrt_amd import RayBatch

    rb = RayBatch.from_numpy(o, d, W)
    # schedule -1: the grid with its blocks dispatched last first; chunk: the persistent
    # schedule's chunk or the compact schedule's first-pass cap
    h = world.cast(rb, max_iters, with_pos=with_pos, schedule=abs(schedule), chunk=chunk if schedule == 2 else 0,
                   counters=counters, order=1 if schedule < 0 else 0, pass_cap=chunk if schedule == 3 else 0)
    torch.cuda.synchronize()
    return h.t.cpu().numpy(), h.rec.cpu().numpy(), (h.pos.cpu().numpy() if h.pos is not None else None)


def pair_from_dense(vrt, oracle_mod, vox):
    sc = vrt.Scene.from_dense(vox)
    return vrt.World.from_scene(sc), dense_world(oracle_mod, vox)


def test_golden_fixture_on_gpu(vrt, oracle_mod, cuda):
    f = np.load(GOLDEN / "blob64.npz")
    world, _ = pair_from_dense(vrt, oracle_mod, f["voxels"])
    for sched, chunk in SCHEDULES:
        t, rec, _ = gpu_cast(world, f["o"], f["d"], tuple(int(x) for x in f["world_origin"]), int(f["max_iters"]),
                             sched, chunk)
        assert_hits_equal(t, rec, f["t"], f["rec"], f"golden sched={sched} chunk={chunk}")
        assert np.array_equal(t.view(np.uint32), f["t"].view(np.uint32))


def test_c1_terrain_parity(vrt, oracle_mod, cuda):
    from voxelrt_amd import workloads as W
    import oracle

    sc = W.C1_WORLD.build()
    world = vrt.World.from_scene(sc)
    ow = oracle.OracleWorld.from_stream(sc.stream())
    cam = W.c1_camera()
    o, d = cam.rays_host()
    ref = ow.cast(o, d, cam.world_origin, W.MAX_ITERS, with_pos=True)
    for sched, chunk in SCHEDULES:
        t, rec, pos = gpu_cast(world, o, d, cam.world_origin, W.MAX_ITERS, sched, chunk, with_pos=True)
        assert_hits_equal(t, rec, ref["t"], ref["rec"], f"C1 sched={sched}")
        assert np.array_equal(t.view(np.uint32), ref["t"].view(np.uint32)), "t not bit-exact"
        hit = (rec[:, 3] & 0xFF) != 0
        assert np.array_equal(pos[:, hit].view(np.uint32), ref["pos"][:, hit].view(np.uint32))
    assert ((ref["rec"][:, 3] & 0xFF) != 0).mean() > 0.2


@pytest.mark.parametrize("seed,size,inside,max_iters", [(1, 64, True, 128), (2, 64, False, 128),
                                                        (3, 128, True, 512), (4, 128, False, 40)])
def test_random_world_parity(vrt, oracle_mod, cuda, seed, size, inside, max_iters):
    vox = blob_world(size, seed=seed, density=0.01)
    world, ow = pair_from_dense(vrt, oracle_mod, vox)
    o, d = random_rays(30000, size, seed=seed + 50, inside=inside)
    # world origins of each specialization of the step: zero, any, a sector corner
    W = {1: (0, 0, 0), 2: (-7, 3, 11), 3: (32, -64, 96), 4: (-32, 0, 64)}[seed]
    o = (o - np.array(W, np.float32)[:, None]).astype(np.float32)
    ref = ow.cast(o, d, W, max_iters)
    for sched, chunk in SCHEDULES:
        t, rec, _ = gpu_cast(world, o, d, W, max_iters, sched, chunk)
        assert_hits_equal(t, rec, ref["t"], ref["rec"], f"seed={seed} sched={sched}")
        assert np.array_equal(t.view(np.uint32), ref["t"].view(np.uint32))


@pytest.mark.parametrize("size,ysize,seed", [(128, 64, 21), (64, 128, 22), (128, 32, 23)])
def test_non_cubic_world_parity(vrt, oracle_mod, cuda, size, ysize, seed):
    """Worlds whose y extent differs from x/z take the two-compare bounds test; rays start
    inside and around the box so that every face of it ends some of them."""
    vox = blob_world(size, seed=seed, density=0.01, ysize=ysize)
    world, ow = pair_from_dense(vrt, oracle_mod, vox)
    info = world.info()
    assert (info.size_xz, info.size_y) == (size, ysize)
    for inside in (True, False):
        o, d = random_rays(20000, size, seed=seed + 100, ysize=ysize, inside=inside)
        ref = ow.cast(o, d, (0, 0, 0), 128)
        for sched, chunk in SCHEDULES:
            t, rec, _ = gpu_cast(world, o, d, (0, 0, 0), 128, sched, chunk)
            assert_hits_equal(t, rec, ref["t"], ref["rec"], f"{size}x{ysize} inside={inside} sched={sched}")
            assert np.array_equal(t.view(np.uint32), ref["t"].view(np.uint32))
        assert ((ref["rec"][:, 3] & 0xFF) != 0).mean() > 0.05


@pytest.mark.parametrize("max_iters", [1, 2, 7])
def test_tiny_iteration_caps(vrt, oracle_mod, cuda, max_iters):
    """Caps of a few lookups, every schedule (a cap of 0 is rejected: the hit record counts
    lookups in 16 bits and a cast takes 1..65535)."""
    vox = blob_world(64, seed=31, density=0.05)
    world, ow = pair_from_dense(vrt, oracle_mod, vox)
    o, d = random_rays(5000, 64, seed=32)
    ref = ow.cast(o, d, (0, 0, 0), max_iters)
    for sched, chunk in SCHEDULES:
        t, rec, _ = gpu_cast(world, o, d, (0, 0, 0), max_iters, sched, chunk)
        assert_hits_equal(t, rec, ref["t"], ref["rec"], f"cap={max_iters} sched={sched} chunk={chunk}")
    if max_iters == 1:
        with pytest.raises(vrt.VrtError, match="max_iters"):
            gpu_cast(world, o, d, (0, 0, 0), 0, 5, 0)


def test_reverse_block_order(vrt, oracle_mod, cuda):
    """The grid's reverse block order writes the same hits (also with a device ray count and a
    partial last block)."""
    import voxelrt_amd as V
    from voxelrt_amd import RayBatch

    vox = blob_world(128, seed=41, density=0.01)
    world, ow = pair_from_dense(vrt, oracle_mod, vox)
    o, d = random_rays(10001, 128, seed=42)
    ref = ow.cast(o, d, (0, 0, 0), 128, with_pos=True)
    rb = RayBatch.from_numpy(o, d, (0, 0, 0))
    for order in (V.VRT_ORDER_FORWARD, V.VRT_ORDER_REVERSE):
        h = world.cast(rb, 128, with_pos=True, order=order)
        torch.cuda.synchronize()
        assert_hits_equal(h.t.cpu().numpy(), h.rec.cpu().numpy(), ref["t"], ref["rec"], f"order={order}")
        hit = (ref["rec"][:, 3] & 0xFF) != 0
        assert np.array_equal(h.pos.cpu().numpy()[:, hit].view(np.uint32), ref["pos"][:, hit].view(np.uint32))
        cnt = torch.tensor([7777], dtype=torch.int32, device="cuda")
        h2 = V.HitBatch.empty(rb.n)
        h2.rec.fill_(-7)
        world.cast(rb, 128, out=h2, count=cnt, order=order)
        torch.cuda.synchronize()
        rec2 = h2.rec.cpu().numpy()
        assert np.array_equal(rec2[:7777], ref["rec"][:7777]) and np.all(rec2[7777:] == -7)
    with pytest.raises(V.VrtError, match="order"):
        world.cast(rb, 128, order=2)


def test_special_rays(vrt, oracle_mod, cuda):
    """Exact zeros (inf/NaN side distances), denormal and huge components, far origins (floor2i
    out of range), origins on voxel faces, empty and solid worlds."""
    vox = blob_world(64, seed=9, density=0.02)
    full = np.full((64, 64, 64), 5, np.uint8)
    empty = np.zeros((64, 64, 64), np.uint8)
    rng = np.random.default_rng(1)
    n = 4096
    o = rng.uniform(-4, 68, (3, n)).astype(np.float32)
    o[:, :512] = np.floor(o[:, :512])                       # on voxel corners
    o[0, 512:600] = np.float32(3e9)                         # floor2i out of range
    o[1, 600:700] = np.float32(-3e9)
    o[2, 700:720] = np.float32(np.nan)
    d = rng.normal(size=(3, n)).astype(np.float32)
    d[1, 1000:1500] = 0.0                                   # exact zeros
    d[0, 1500:1800] = 0.0
    d[2, 1500:1800] = 0.0
    d[0, 1800:2000] = np.float32(1e-40)                     # denormal
    d[2, 2000:2100] = np.float32(-1e-41)
    d[1, 2100:2200] = np.float32(1e30)                      # huge
    d[:, 2200:2300] = np.float32(-0.0)
    d[0, 2200:2300] = 1.0
    for v in (vox, full, empty):
        world, ow = pair_from_dense(vrt, oracle_mod, v)
        ref = ow.cast(o, d, (0, 0, 0), 128)
        for sched, chunk in SCHEDULES:
            t, rec, _ = gpu_cast(world, o, d, (0, 0, 0), 128, sched, chunk)
            assert_hits_equal(t, rec, ref["t"], ref["rec"], f"special sched={sched}")
            assert np.array_equal(t.view(np.uint32), ref["t"].view(np.uint32))


def test_counters_match_oracle(vrt, oracle_mod, cuda):
    vox = blob_world(128, seed=12, density=0.01)
    world, ow = pair_from_dense(vrt, oracle_mod, vox)
    o, d = random_rays(20000, 128, seed=13)
    ref = ow.count(o, d, (0, 0, 0), 128)
    for sched, chunk in SCHEDULES:
        ctr = torch.zeros(4, dtype=torch.int64, device="cuda")
        gpu_cast(world, o, d, (0, 0, 0), 128, sched, chunk, counters=ctr)
        c = ctr.cpu().numpy()
        assert (c[0], c[1], c[2], c[3]) == (ref["rays"], ref["iters"], ref["descents"], ref["hits"])


def test_device_camera_rays_equal_host(vrt, cuda):
    from voxelrt_amd import workloads as W

    for tile in [(1, 1), (8, 8), (16, 4)]:
        cam = W.c2_camera(tile=tile, width=640, height=360)
        ho, hd = cam.rays_host()
        rb = cam.rays_device()
        torch.cuda.synchronize()
        assert np.array_equal(rb.o.cpu().numpy().view(np.uint32), ho.view(np.uint32))
        assert np.array_equal(rb.d.cpu().numpy().view(np.uint32), hd.view(np.uint32))


def test_fp64_picking_cast_matches_oracle(vrt, oracle_mod, cuda):
    vox = blob_world(64, seed=7, density=0.004, blobs=3)
    world, ow = pair_from_dense(vrt, oracle_mod, vox)
    o, d = random_rays(300, 64, seed=78, inside=False)
    for i in range(o.shape[1]):
        oo, dd = o[:, i].astype(np.float64), d[:, i].astype(np.float64)
        g = world.ray_cast(oo, dd, 1024)
        r = ow.cast_f64(oo, dd, 1024)
        assert g.distance == r["distance"] and g.voxel_pos == r["voxel"] and g.normal == r["normal"]
        assert g.uv == r["uv"]


@pytest.mark.parametrize("anchor", ["floor", "sector", "zero"])
def test_counting_variant_writes_the_same_hits(vrt, cuda, anchor):
    """The cast kernel's counting instantiation (the one that fills vrt_cast_counters) traces
    the same rays to the same records as the plain one, on camera and bounce rays of each
    world-origin specialization; its counts agree with the records."""
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world = V.World.from_scene(W.C1_WORLD.build())
    cam = V.CameraSpec((128.37, 200.61, 12.29), yaw_deg=8.0, pitch_deg=-28.0, width=640, height=360, tile=(8, 8),
                       anchor=anchor)
    prim = cam.rays_device()
    hp = world.cast(prim, W.MAX_ITERS, with_pos=True)
    br, _, cnt = V.bounce(prim, hp, cam.c.width, 9)
    torch.cuda.synchronize()
    nb = int(cnt.item())
    for name, rays, n, kw in [("primary", prim, prim.n, {}), ("bounce", br, nb, {"count": cnt})]:
        plain = world.cast(rays, W.MAX_ITERS, **kw)
        ctr = V.new_counters()
        counted = world.cast(rays, W.MAX_ITERS, counters=ctr, **kw)
        torch.cuda.synchronize()
        assert torch.equal(plain.rec[:n], counted.rec[:n]), f"{anchor} {name}"
        assert torch.equal(plain.t[:n], counted.t[:n]), f"{anchor} {name}"
        c = ctr.cpu().numpy()
        rec = plain.rec[:n].cpu().numpy()
        assert c[0] == n and c[3] == int(((rec[:, 3] & 0xFF) != 0).sum()), f"{anchor} {name}"
        assert c[1] == int(((rec[:, 3] >> 16) & 0xFFFF).sum()), f"{anchor} {name}"


@pytest.mark.parametrize("W", [(0, 0, 0), (32, -64, 96), (-7, 3, 11)])
def test_counting_variant_non_cubic_world(vrt, oracle_mod, cuda, W):
    """The counting instantiation on a non-cubic world (the two-compare bounds test after the
    loop) at each world-origin specialization: records bit-exact with the oracle, counter
    totals equal to the oracle's."""
    vox = blob_world(128, seed=41, density=0.01, ysize=64)
    world, ow = pair_from_dense(vrt, oracle_mod, vox)
    for inside in (True, False):
        o, d = random_rays(20000, 128, seed=42, ysize=64, inside=inside)
        o = (o - np.array(W, np.float32)[:, None]).astype(np.float32)
        ref = ow.cast(o, d, W, 128)
        cref = ow.count(o, d, W, 128)
        for sched in (1, -1):
            ctr = vrt.new_counters()
            t, rec, _ = gpu_cast(world, o, d, W, 128, sched, counters=ctr)
            assert_hits_equal(t, rec, ref["t"], ref["rec"], f"W={W} inside={inside} sched={sched}")
            c = ctr.cpu().numpy()
            assert (c[0], c[1], c[2], c[3]) == (o.shape[1], cref["iters"], cref["descents"], cref["hits"])


@pytest.mark.parametrize("W", [(0, 0, 0), (32, 64, 96), (5, -3, 7)])
def test_uv_and_material_data_match_oracle(vrt, oracle_mod, cuda, W):
    """The rest of VHitResult (CpuRenderer.cpp:210-221): UV bit-exact with the oracle's, and the
    palette entry's low 32 bits of the hit material, for every world-origin specialization and
    every schedule (the compact one writes its survivors' fields in its second pass); other
    entry points reject the fields."""
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L

    vox = blob_world(128, seed=51, density=0.01)
    world, ow = pair_from_dense(vrt, oracle_mod, vox)
    rng = np.random.default_rng(52)
    palette = rng.integers(0, 2**63, 256, dtype=np.int64).astype(np.uint64)
    world.set_palette(palette)
    o, d = random_rays(20000, 128, seed=53)
    o = (o - np.array(W, np.float32)[:, None]).astype(np.float32)
    ref = ow.cast(o, d, W, 128, with_pos=True, with_uv=True)
    rb = V.RayBatch.from_numpy(o, d, W)
    for sched, chunk in [(1, 0), (2, 64), (3, 0), (3, 5)]:
        h = world.cast(rb, 128, with_pos=True, schedule=sched, chunk=chunk if sched == 2 else 0,
                       pass_cap=chunk if sched == 3 else 0, with_uv=True, with_material=True)
        torch.cuda.synchronize()
        assert_hits_equal(h.t.cpu().numpy(), h.rec.cpu().numpy(), ref["t"], ref["rec"], f"W={W} sched={sched}")
        assert np.array_equal(h.uv.cpu().numpy().view(np.uint32), ref["uv"].view(np.uint32))
        mat = (ref["rec"][:, 3] & 0xFF).astype(np.int64)
        want = np.where(mat != 0, (palette[mat] & np.uint64(0xFFFFFFFF)).astype(np.uint32), 0).astype(np.uint32)
        assert np.array_equal(h.material_data.cpu().numpy().view(np.uint32), want)
    hit = (ref["rec"][:, 3] & 0xFF) != 0
    assert hit.mean() > 0.2 and np.all(ref["uv"][:, hit] >= 0) and np.all(ref["uv"][:, hit] <= 1)
    with pytest.raises(V.VrtError, match="filled by vrt_cast"):
        world.cast_shader(rb, out=V.HitBatch.empty(rb.n, with_uv=True))
