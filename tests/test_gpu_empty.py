"""Zero-size inputs through every entry point that takes a count: the call succeeds, writes
nothing it should not, and option errors are reported whatever the count."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_world():
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    scene = W.C1_WORLD.build()
    return V.World.from_scene(scene), W.c1_camera()


def test_zero_rays_every_schedule(vrt, cuda, small_world):
    import voxelrt_amd as V

    world, _ = small_world
    empty = V.RayBatch.empty(0)
    for sched, chunk in [(0, 0), (1, 0), (2, 64), (3, 0)]:
        ctr = V.new_counters()
        h = world.cast(empty, 128, with_pos=True, schedule=sched, chunk=chunk, counters=ctr)
        torch.cuda.synchronize()
        assert h.t.numel() == 0 and int(ctr.sum().item()) == 0
    for bad in (dict(order=2), dict(schedule=9), dict(schedule=4), dict(schedule=5), dict(schedule=2, chunk=100)):
        with pytest.raises(V.VrtError):
            world.cast(empty, 128, **bad)
    with pytest.raises(V.VrtError):
        world.cast(empty, 0)


def test_zero_rays_bounce_generation_writes_zero_count(vrt, cuda, small_world):
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L

    empty = V.RayBatch.empty(0)
    count = torch.full((1,), 123, dtype=torch.int32, device=cuda)
    src = torch.empty(1, dtype=torch.int32, device=cuda)
    scratch = torch.empty(int(L.lib().vrt_bounce_scratch_bytes(0)) // 4 + 1, dtype=torch.int32, device=cuda)
    # a 1-element source buffer: the entry point wants non-null outputs even for no rays
    obuf = torch.empty(6, dtype=torch.float32, device=cuda)
    bufs = L.RayBuffers(*[obuf[i].data_ptr() for i in range(6)])
    hbuf = torch.empty(16, dtype=torch.int32, device=cuda)
    pbuf = torch.empty(3, dtype=torch.float32, device=cuda)
    h = L.Hits(pbuf.data_ptr(), hbuf.data_ptr(), pbuf.data_ptr())
    r = empty.to_c()
    L.check(L.lib().vrt_bounce_generate(C.byref(r), C.byref(h), 64, 1, C.byref(bufs), src.data_ptr(),
                                        count.data_ptr(), scratch.data_ptr(), None))
    torch.cuda.synchronize()
    assert int(count.item()) == 0
    # with nothing to generate, only the count pointer is needed
    count.fill_(55)
    nul = L.RayBuffers(0, 0, 0, 0, 0, 0)
    L.check(L.lib().vrt_bounce_generate(C.byref(r), C.byref(L.Hits(0, 0, 0)), 64, 1, C.byref(nul), None,
                                        count.data_ptr(), None, None))
    torch.cuda.synchronize()
    assert int(count.item()) == 0
    # and a cast given that device count of 0 traces nothing
    prim = small_world[1].rays_device()
    hb = V.HitBatch.empty(prim.n)
    hb.rec.fill_(-7)
    world = small_world[0]
    world.cast(prim, 128, out=hb, count=count)
    torch.cuda.synchronize()
    assert bool(torch.all(hb.rec == -7))


def test_batch_with_empty_segments(vrt, cuda, small_world):
    import voxelrt_amd as V

    world, cam = small_world
    o, d = cam.rays_host()
    rays = V.RayBatch.from_numpy(o[:, :5000], d[:, :5000], cam.world_origin)
    full = V.Batch([cam], [(0, 5000, cam.world_origin, None)], rays=rays)
    V.BatchPlan(world, full, (8, 8, 1, 0)).launch()
    # the same segments with empty ray segments before, between and after them
    holey = V.Batch([cam], [(0, 0, (0, 0, 0), None), (0, 5000, cam.world_origin, None),
                            (100, 0, (0, 0, 0), None)], rays=rays)
    for tuning in [(8, 8, 1, 0), (16, 4, 2, 64)]:
        holey.hits.rec.fill_(-7)
        V.BatchPlan(world, holey, tuning).launch()
        torch.cuda.synchronize()
        for k_full, k_holey in [(0, 0), (1, 2)]:
            t_f, r_f, _ = full.segment_hits(k_full)
            t_h, r_h, _ = holey.segment_hits(k_holey)
            assert torch.equal(r_f, r_h) and torch.equal(t_f, t_h)
        assert holey.segment_hits(1)[1].numel() == 0 and holey.segment_hits(3)[1].numel() == 0
    # a batch with nothing to cast at all
    none = V.Batch([], [(0, 0, (0, 0, 0), None)], rays=None)
    plan = V.BatchPlan(world, none, (8, 8, 1, 0))
    plan.launch()
    plan.graph_launch()
    torch.cuda.synchronize()


def test_render_with_no_views_and_no_pixels(vrt, cuda, small_world):
    import voxelrt_amd as V

    world, _ = small_world
    nvox = world.info().num_bricks * 512
    albedo = torch.rand((nvox, 3), device=cuda)
    img, sv, sw = V.render_saved(world, albedo, [])
    torch.cuda.synchronize()
    assert img.numel() == 0 and sv.numel() == 0 and sw.numel() == 0
    g = V.render_backward(world, torch.empty((0, 3), device=cuda), torch.empty(0, dtype=torch.int32, device=cuda),
                          torch.empty(0, device=cuda), nvox)
    torch.cuda.synchronize()
    assert g.shape == (nvox, 3) and not bool(g.any())


def test_render_rejects_albedo_shorter_than_the_slots(vrt, cuda, small_world):
    import voxelrt_amd as V

    world, cam = small_world
    info = world.info()
    assert info.slot_high_water == info.num_bricks  # a fresh world fills slots 0 .. num_bricks - 1
    nvox = info.slot_high_water * 512
    albedo = torch.rand((nvox - 1, 3), device=cuda)
    with pytest.raises(RuntimeError, match="slot_high_water"):
        V.render(world, albedo, [cam])
    with pytest.raises(RuntimeError, match="slot_high_water"):
        V.render_saved(world, albedo, [cam])
    with pytest.raises(ValueError):
        V.render_backward(world, torch.zeros((4, 3), device=cuda), torch.zeros(4, dtype=torch.int32, device=cuda),
                          torch.zeros(4, device=cuda), nvox - 1)
