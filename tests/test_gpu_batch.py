"""GPU checks of the batched path: one launch over uneven camera and ray segments equals
plain per-segment casts (and the oracle), for every tuning, direct and hipGraph-replayed."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import assert_hits_equal

pytestmark = pytest.mark.gpu

TUNINGS = [(8, 8, 1, 0), (16, 4, 1, 0), (8, 8, 2, 0), (16, 4, 2, 0), (8, 8, 1, 64), (16, 4, 2, 256), (4, 16, 1, 128)]


@pytest.fixture(scope="module")
def c2_small():
    """The 1024^3 world with 6 uneven cameras and their bounce sets (truncated to seeded fractions)."""
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    scene = W.C2_WORLD.build()
    world = V.World.from_scene(scene)
    cams = W.c4_cameras(n=6, seed=11, tile=(1, 1))
    rng = np.random.default_rng(12)
    # primary pass per camera (plain cast of materialised rays) -> bounce sets
    prim = [c.rays_device() for c in cams]
    ray_parts, segs, plain = [], [], []
    off = 0
    for c, rb in zip(cams, prim):
        h = world.cast(rb, W.MAX_ITERS, with_pos=True)
        br, src, cnt = V.bounce(rb, h, c.c.width, W.BOUNCE_SEED)
        torch.cuda.synchronize()
        nb = int(int(cnt.item()) * rng.uniform(0.25, 1.0))
        ray_parts.append((br.o[:, :nb], br.d[:, :nb]))
        segs.append((off, nb, rb.world_origin, None))
        off += nb
        plain.append(h)
    rays = V.RayBatch(torch.cat([p[0] for p in ray_parts], 1).contiguous(),
                      torch.cat([p[1] for p in ray_parts], 1).contiguous(), (0, 0, 0))
    return world, scene, cams, prim, segs, rays


def plain_results(world, cams, prim, segs, rays, max_iters):
    import voxelrt_amd as V

    out = []
    for rb in prim:
        h = world.cast(rb, max_iters)
        out.append((h.t.cpu().numpy(), h.rec.cpu().numpy()))
    for (ro, n, W_, _) in segs:
        rb = V.RayBatch(rays.o[:, ro:ro + n].contiguous(), rays.d[:, ro:ro + n].contiguous(), W_)
        h = world.cast(rb, max_iters)
        out.append((h.t.cpu().numpy(), h.rec.cpu().numpy()))
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("tuning", TUNINGS)
def test_batch_equals_plain_casts(vrt, cuda, c2_small, tuning):
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene, cams, prim, segs, rays = c2_small
    ref = plain_results(world, cams, prim, segs, rays, W.MAX_ITERS)
    batch = V.Batch(cams, segs, rays, W.MAX_ITERS)
    plan = V.BatchPlan(world, batch, tuning)
    ctr = V.new_counters()
    plan.launch(counters=ctr)
    torch.cuda.synchronize()
    for k in range(len(ref)):
        t, rec, _ = batch.segment_hits(k)
        assert_hits_equal(t.cpu().numpy(), rec.cpu().numpy(), ref[k][0], ref[k][1], f"tuning {tuning} segment {k}")
    assert int(ctr[0].item()) == batch.total
    plan.close()


def test_invalid_tuning_is_rejected(vrt, cuda, c2_small):
    import voxelrt_amd as V

    world, scene, cams, prim, segs, rays = c2_small
    batch = V.Batch(cams, segs, rays, 128)
    for bad in [(8, 4, 1, 0), (8, 8, 3, 0), (8, 8, 1, 100)]:
        with pytest.raises(V.VrtError):
            V.BatchPlan(world, batch, bad)
    if any(c.c.width % 64 for c in cams):
        with pytest.raises(V.VrtError):
            V.BatchPlan(world, batch, (64, 1, 1, 0))


def test_batch_matches_oracle(vrt, oracle_mod, cuda, c2_small):
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene, cams, prim, segs, rays = c2_small
    ow = oracle_mod.OracleWorld.from_stream(scene.stream())
    batch = V.Batch(cams, segs, rays, W.MAX_ITERS)
    plan = V.BatchPlan(world, batch, (16, 4, 2, 0))
    plan.launch()
    torch.cuda.synchronize()
    for k, rb in enumerate(prim):
        r = ow.cast_simd(rb.o.cpu().numpy(), rb.d.cpu().numpy(), rb.world_origin, W.MAX_ITERS)
        t, rec, _ = batch.segment_hits(k)
        assert_hits_equal(t.cpu().numpy(), rec.cpu().numpy(), r["t"], r["rec"], f"camera {k}")
    for j, (ro, n, W_, _) in enumerate(segs):
        r = ow.cast_simd(rays.o[:, ro:ro + n].cpu().numpy(), rays.d[:, ro:ro + n].cpu().numpy(), W_, W.MAX_ITERS)
        t, rec, _ = batch.segment_hits(len(prim) + j)
        assert_hits_equal(t.cpu().numpy(), rec.cpu().numpy(), r["t"], r["rec"], f"bounce set {j}")


def test_graph_replays_match_eager(vrt, cuda, c2_small):
    """The captured graph (queue reset by a fill kernel, never a memset) gives the eager result on
    every replay, with eager work and synchronisations in between."""
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene, cams, prim, segs, rays = c2_small
    for tuning in [(8, 8, 1, 0), (16, 4, 1, 64)]:
        batch = V.Batch(cams, segs, rays, W.MAX_ITERS)
        plan = V.BatchPlan(world, batch, tuning)
        plan.launch()
        torch.cuda.synchronize()
        t0, r0 = batch.hits.t.clone(), batch.hits.rec.clone()
        for rep in range(3):
            batch.hits.rec.fill_(-3)
            batch.hits.t.fill_(7.0)
            torch.cuda.synchronize()
            plan.graph_launch()
            torch.cuda.synchronize()
            assert torch.equal(batch.hits.rec, r0), f"replay {rep}"
            assert torch.equal(batch.hits.t, t0), f"replay {rep}"
            world.cast(prim[0], W.MAX_ITERS)  # eager work between replays
            torch.cuda.synchronize()
        plan.close()


def test_device_count_segments_and_positions(vrt, cuda, c2_small):
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene, cams, prim, segs, rays = c2_small
    ro, n, W_, _ = segs[0]
    cnt = torch.tensor([n // 3], dtype=torch.int32, device="cuda")
    batch = V.Batch(cams[:1], [(ro, n, W_, cnt)], rays, W.MAX_ITERS, with_pos=True)
    plan = V.BatchPlan(world, batch, (8, 8, 1, 0))
    batch.hits.rec.fill_(-9)
    plan.launch()
    torch.cuda.synchronize()
    t, rec, pos = batch.segment_hits(1)
    rec = rec.cpu().numpy()
    assert np.all(rec[n // 3:] == -9)
    ref = world.cast(V.RayBatch(rays.o[:, ro:ro + n // 3].contiguous(), rays.d[:, ro:ro + n // 3].contiguous(), W_),
                     W.MAX_ITERS, with_pos=True)
    torch.cuda.synchronize()
    assert np.array_equal(rec[:n // 3], ref.rec.cpu().numpy())
    assert torch.equal(pos[:, :n // 3], ref.pos)
    tc, rc, pc = batch.segment_hits(0)
    refc = world.cast(prim[0], W.MAX_ITERS, with_pos=True)
    torch.cuda.synchronize()
    assert torch.equal(rc, refc.rec) and torch.equal(pc, refc.pos)


def test_autotune_picks_a_valid_tuning_and_caches(vrt, cuda, c2_small, tmp_path):
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene, cams, prim, segs, rays = c2_small
    batch = V.Batch(cams, segs, rays, W.MAX_ITERS)
    cache = tmp_path / "tune.txt"
    best, ms = V.autotune(world, batch, str(cache))
    assert best[:2] in [(8, 8), (16, 4)] and best[2] in (1, 2) and best[3] in (0, 64, 256)
    assert ms > 0 and cache.exists() and len(cache.read_text().splitlines()) == 1
    best2, ms2 = V.autotune(world, batch, str(cache))
    assert best2 == best and ms2 == pytest.approx(ms)


def test_segments_with_mixed_world_origins(vrt, cuda, c2_small):
    """Segments whose world origins take the zero, general and aligned specializations of the
    traversal, side by side in one launch, each equal to a plain cast of the same rays."""
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, _, cams, prim, segs, rays = c2_small
    origins = [(0, 0, 0), (-7, 3, 11), (32, -64, 96)]
    parts, msegs = [], []
    off = 0
    for j, (ro, n, W_, _) in enumerate(segs):
        Wn = origins[j % len(origins)]
        shift = torch.tensor([W_[a] - Wn[a] for a in range(3)], dtype=torch.float32, device="cuda")[:, None]
        parts.append(rays.o[:, ro:ro + n] + shift)
        msegs.append((off, n, Wn, None))
        off += n
    mrays = V.RayBatch(torch.cat(parts, 1).contiguous(), rays.d.contiguous(), (0, 0, 0))
    ref = plain_results(world, cams, prim, msegs, mrays, W.MAX_ITERS)
    for tuning in ((16, 4, 1, 0), (8, 8, 2, 0)):
        batch = V.Batch(cams, msegs, mrays, W.MAX_ITERS)
        V.BatchPlan(world, batch, tuning).launch()
        torch.cuda.synchronize()
        for k in range(len(prim) + len(msegs)):
            t, rec, _ = batch.segment_hits(k)
            assert np.array_equal(rec.cpu().numpy(), ref[k][1]), f"segment {k} tuning {tuning}"
            assert np.array_equal(t.cpu().numpy().view(np.uint32), ref[k][0].view(np.uint32)), f"segment {k}"


def test_graph_recaptured_after_slot_growth(vrt, oracle_mod, cuda):
    """A sync that outgrows the slot buffers moves them (the world's generation changes): the next
    graph replay captures again with the new buffers and equals an eager launch and the oracle."""
    import voxelrt_amd as V
    from test_gpu_world import sector_stream_of
    from test_oracle import dense_world

    vox = np.zeros((128, 128, 128), np.uint8)
    vox[60:70, 40:50, 50:60] = 3  # a few bricks: the slot buffers start at their minimum, 1024 slots
    world = V.World(7, 7)
    world.sync(*sector_stream_of(vox, list(range(64))))
    i0 = world.info()
    assert i0.slot_capacity == 1024
    cams = [V.CameraSpec((64.5, 100.25, 10.75), 20.0, -35.0, 64, 64),
            V.CameraSpec((20.5, 90.5, 100.5), 200.0, -50.0, 128, 64, anchor="sector")]
    batch = V.Batch(cams, max_iters=128)
    plan = V.BatchPlan(world, batch, (8, 8, 1, 0))
    plan.graph_launch()
    torch.cuda.synchronize()
    # a solid floor 48 voxels deep: 16 x 16 x 6 = 1536 more bricks than the slot buffers hold
    vox2 = vox.copy()
    vox2[:48] = 4
    world.sync(*sector_stream_of(vox2, list(range(64))))
    i1 = world.info()
    assert i1.slot_capacity > i0.slot_capacity and i1.generation != i0.generation
    ow = dense_world(oracle_mod, vox2)
    for rep in range(2):
        batch.hits.rec.fill_(-3)
        torch.cuda.synchronize()
        plan.graph_launch()
        torch.cuda.synchronize()
        g_t, g_rec = batch.hits.t.clone(), batch.hits.rec.clone()
        plan.launch()
        torch.cuda.synchronize()
        assert torch.equal(g_rec, batch.hits.rec) and torch.equal(g_t, batch.hits.t), f"replay {rep}"
        for k, c in enumerate(cams):
            o, d = c.rays_host()
            ref = ow.cast(o, d, c.world_origin, 128)
            t, rec, _ = batch.segment_hits(k)
            assert_hits_equal(t.cpu().numpy(), rec.cpu().numpy(), ref["t"], ref["rec"], f"camera {k} replay {rep}")
    plan.close()


def test_plan_rejects_32bit_index_overflow(vrt, cuda, c2_small):
    """Hit and ray indices are 32-bit in the kernel: segments reaching past index 2^32 - 1 and
    cameras of more than 2^32 - 256 pixels are rejected when the plan is built."""
    import ctypes as C

    import voxelrt_amd as V
    from voxelrt_amd import _lib as L

    world, scene, cams, prim, segs, rays = c2_small
    for hit_off, ray_off, cnt in [(0xFFFFFF00, 0, 0x200), (0, 0xFFFFFFF0, 64), (0, 0, 0xFFFFFFF0)]:
        seg = L.Segment(L.VRT_SEG_RAYS, 0, ray_off, hit_off, cnt, (C.c_int32 * 3)(0, 0, 0), 0)
        b = L.Batch(None, 0, C.cast(C.pointer(seg), C.c_void_p), 1, rays.to_c(), V.HitBatch.empty(64).to_c(), 128)
        out = C.c_void_p()
        t = L.BatchTuning(8, 8, 1, 0)
        assert V.lib().vrt_batch_plan_create(world.handle, C.byref(b), C.byref(t), None, C.byref(out)) == \
            L.VRT_ERR_INVALID_ARGUMENT
        assert b"2^32" in V.lib().vrt_last_error()
    big = L.Camera.from_buffer_copy(cams[0].c)
    big.width, big.height, big.tile_w, big.tile_h = 65536, 65536, 1, 1
    seg = L.Segment(L.VRT_SEG_CAMERA, 0, 0, 0, 0, (C.c_int32 * 3)(0, 0, 0), 0)
    b = L.Batch(C.cast(C.pointer(big), C.c_void_p), 1, C.cast(C.pointer(seg), C.c_void_p), 1, L.Rays(),
                V.HitBatch.empty(64).to_c(), 128)
    out = C.c_void_p()
    t = L.BatchTuning(8, 8, 1, 0)
    assert V.lib().vrt_batch_plan_create(world.handle, C.byref(b), C.byref(t), None, C.byref(out)) == \
        L.VRT_ERR_INVALID_ARGUMENT
    assert b"2^32 - 256 pixels" in V.lib().vrt_last_error()


def test_c4_full_batch_matches_oracle(vrt, oracle_mod, cuda, c2_small):
    """The whole batched configuration: 64 cameras of uneven square sizes (128..1024 px) and each
    camera's bounce set truncated to a seeded fraction U(0.25, 1), 128 segments traced in one
    hipGraph launch, compared per pixel and per bounce ray with the oracle."""
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    world, scene = c2_small[0], c2_small[1]
    cams = W.c4_cameras(64, seed=4, tile=(1, 1))
    rng = np.random.default_rng(44)
    parts, segs = [], []
    off = 0
    for c in cams:
        rb = c.rays_device()
        h = world.cast(rb, W.MAX_ITERS, with_pos=True)
        br, _, cnt = V.bounce(rb, h, c.c.width, W.BOUNCE_SEED)
        torch.cuda.synchronize()
        nb = int(int(cnt.item()) * rng.uniform(0.25, 1.0))
        parts.append((br.o[:, :nb].clone(), br.d[:, :nb].clone()))
        segs.append((off, nb, rb.world_origin, None))
        off += nb
        del rb, h, br
    rays = V.RayBatch(torch.cat([p[0] for p in parts], 1).contiguous(),
                      torch.cat([p[1] for p in parts], 1).contiguous(), (0, 0, 0))
    del parts
    batch = V.Batch(cams, segs, rays, W.MAX_ITERS)
    plan = V.BatchPlan(world, batch, (16, 4, 1, 0))
    plan.graph_launch()
    torch.cuda.synchronize()
    n_prim = sum(c.n for c in cams)
    print(f"C4: {n_prim} primary + {off} bounce rays in {len(cams) + len(segs)} segments", flush=True)
    assert n_prim > 20_000_000
    ow = oracle_mod.OracleWorld.from_stream(scene.stream())
    for k, c in enumerate(cams):
        o, d = c.rays_host()
        r = ow.cast_simd(o, d, c.world_origin, W.MAX_ITERS)
        t, rec, _ = batch.segment_hits(k)
        assert_hits_equal(t.cpu().numpy(), rec.cpu().numpy(), r["t"], r["rec"], f"camera {k}")
    for j, (ro, n, W_, _) in enumerate(segs):
        r = ow.cast_simd(rays.o[:, ro:ro + n].cpu().numpy(), rays.d[:, ro:ro + n].cpu().numpy(), W_, W.MAX_ITERS)
        t, rec, _ = batch.segment_hits(len(cams) + j)
        assert_hits_equal(t.cpu().numpy(), rec.cpu().numpy(), r["t"], r["rec"], f"bounce set {j}")
    plan.close()
