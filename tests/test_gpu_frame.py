"""GPU checks of the fused frame (vrt_cast_frame): primary hits equal vrt_cast on the camera's
rays, the bounce hit of pixel i equals vrt_cast of the bounce ray vrt_bounce_generate builds for
primary hit i, for the grid and the persistent schedule."""
from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

# (schedule, refill, chunk, order): grid in both block orders, persistent with several refills
SCHEDS = [(1, 0, 0, 0), (1, 0, 0, 1), (2, 16, 256, 0), (2, 8, 64, 0), (2, 32, 1024, 0), (2, 64, 128, 0)]


# c2 at each camera anchor: the sector corner, floor(eye) and the world's own origin take the
# aligned, general and zero-origin specializations of the traversal
@pytest.mark.parametrize("config", ["c1", "c2", "c2-floor", "c2-zero"])
def test_frame_equals_separate_pipeline(vrt, cuda, config):
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L
    from voxelrt_amd import workloads as W

    if config == "c1":
        scene, cam = W.C1_WORLD.build(), W.c1_camera()
    else:
        anchor = config.split("-")[1] if "-" in config else "sector"
        scene, cam = W.C2_WORLD.build(), W.c2_camera(anchor=anchor)
    world = V.World.from_scene(scene)
    prim = cam.rays_device()
    hp_ref = world.cast(prim, W.MAX_ITERS, with_pos=True)
    br, src, cnt = V.bounce(prim, hp_ref, cam.c.width, 9)
    hb_ref = world.cast(br, W.MAX_ITERS, count=cnt)
    torch.cuda.synchronize()
    nb = int(cnt.item())
    src = src[:nb].long()
    c_ref = V.new_counters()
    world.cast(prim, W.MAX_ITERS, counters=c_ref)
    world.cast(br, W.MAX_ITERS, count=cnt, counters=c_ref)
    for sched, refill, chunk, order in SCHEDS:
        ctr = V.new_counters()
        hp, hb = V.cast_frame(world, cam, W.MAX_ITERS, 9, with_pos=True, schedule=sched, refill=refill, chunk=chunk,
                              counters=ctr, order=order)
        torch.cuda.synchronize()
        what = f"{config} sched={sched} refill={refill} chunk={chunk} order={order}"
        assert torch.equal(hp.rec, hp_ref.rec), what
        assert torch.equal(hp.t, hp_ref.t), what
        hit = (hp.rec[:, 3] & 0xFF) != 0
        assert torch.equal(hp.pos[:, hit], hp_ref.pos[:, hit]), what
        assert torch.equal(hb.rec[src], hb_ref.rec[:nb]), what
        assert torch.equal(hb.t[src], hb_ref.t[:nb]), what
        miss = ~hit
        assert torch.all(hb.rec[miss, 3] == L.VRT_HIT_NO_RAY) and torch.all(hb.t[miss] == -1.0), what
        assert torch.equal(ctr, c_ref), what
