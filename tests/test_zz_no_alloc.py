"""The no-allocation bounce generation (vrt_bounce_generate_noalloc, through the C ABI) against
the workspace variant (vrt_bounce_generate): identical rays, sources and counts, on two calls in
a row, the second with fewer primaries; sizes beyond its scratch are rejected."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def outputs(V, n):
    rays = V.RayBatch.empty(n)
    src = torch.full((n,), -5, dtype=torch.int32, device="cuda")
    count = torch.full((1,), -5, dtype=torch.int32, device="cuda")
    return rays, src, count


def test_noalloc_bounce_generation_matches_workspace_variant(vrt, cuda):
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L
    from voxelrt_amd import workloads as W

    lib = V.lib()
    world = V.World.from_scene(W.C1_WORLD.build())
    # two calls in a row, the second smaller: 1280x720 then 640x360 primaries
    for k, (w, h) in enumerate([(1280, 720), (640, 360)]):
        cam = W.c1_camera(width=w, height=h)
        prim = cam.rays_device()
        hits = world.cast(prim, W.MAX_ITERS, with_pos=True)
        r, hc = prim.to_c(), hits.to_c()
        ref_rays, ref_src, ref_cnt = outputs(V, prim.n)
        ws = torch.empty(int(lib.vrt_bounce_scratch_bytes(prim.n)) // 4 + 1, dtype=torch.int32, device="cuda")
        L.check(lib.vrt_bounce_generate(C.byref(r), C.byref(hc), w, W.BOUNCE_SEED, C.byref(ref_rays.buffers()),
                                        ref_src.data_ptr(), ref_cnt.data_ptr(), ws.data_ptr(), None))
        na_rays, na_src, na_cnt = outputs(V, prim.n)
        L.check(lib.vrt_bounce_generate_noalloc(C.byref(r), C.byref(hc), w, W.BOUNCE_SEED,
                                                C.byref(na_rays.buffers()), na_src.data_ptr(), na_cnt.data_ptr(),
                                                None))
        torch.cuda.synchronize()
        nb = int(ref_cnt.item())
        assert nb > 1000 and int(na_cnt.item()) == nb, k
        assert torch.equal(na_src[:nb], ref_src[:nb]), k
        assert torch.equal(na_rays.o[:, :nb], ref_rays.o[:, :nb]), k
        assert torch.equal(na_rays.d[:, :nb], ref_rays.d[:, :nb]), k
        # nothing written past the count
        assert bool((na_src[nb:] == -5).all()), k
    # more primaries than the scratch covers: rejected before any launch
    big = L.Rays(0x1000, 0x1000, 0x1000, 0x1000, 0x1000, 0x1000, (L.i32 * 3)(0, 0, 0), 8388609)
    fake_hits = L.Hits(0x1000, 0x1000, 0x1000)
    fake_out = L.RayBuffers(0x1000, 0x1000, 0x1000, 0x1000, 0x1000, 0x1000)
    rc = lib.vrt_bounce_generate_noalloc(C.byref(big), C.byref(fake_hits), 1920, 1, C.byref(fake_out), 0x1000, 0x1000,
                                         None)
    assert rc == L.VRT_ERR_INVALID_ARGUMENT and b"8388608" in lib.vrt_last_error()
    # no primaries: the count is set to zero
    empty = V.RayBatch.empty(0)
    cnt = torch.full((1,), 9, dtype=torch.int32, device="cuda")
    src = torch.empty(1, dtype=torch.int32, device="cuda")
    ob = torch.empty(6, dtype=torch.float32, device="cuda")
    bufs = L.RayBuffers(*[ob.data_ptr() + 4 * i for i in range(6)])
    eh = V.HitBatch.empty(0, with_pos=True)
    L.check(lib.vrt_bounce_generate_noalloc(C.byref(empty.to_c()), C.byref(eh.to_c()), 640, 1, C.byref(bufs),
                                            src.data_ptr(), cnt.data_ptr(), None))
    torch.cuda.synchronize()
    assert int(cnt.item()) == 0
