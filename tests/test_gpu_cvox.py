"""The world-file path end to end on the GPU: the committed "cvox 0004" file (written by the
independent writer of tests/golden/make_cvox.py, layout of VoxelMap.cpp:207-274) loaded by
vrt_scene_load_cvox, uploaded, and cast with a 640x360 camera, against the oracle world built
from the dense grid the file was written from."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

from helpers import assert_hits_equal

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("anchor", ["floor", "sector", "zero"])
def test_world_file_cast_matches_oracle(vrt, oracle_mod, cuda, anchor):
    import voxelrt_amd as V
    from test_oracle import dense_world

    f = np.load(GOLDEN / "terrain128.npz")
    vox = f["voxels"]
    scene, info = V.Scene.load_cvox(str(GOLDEN / "terrain128.cvox"), 7, 7,
                                    sector_offset=tuple(int(x) for x in f["sector_offset"]))
    assert info["sectors_loaded"] == info["sectors_in_file"] == 35 and info["sectors_outside"] == 0
    world = V.World.from_scene(scene)
    world.set_palette(scene.palette())
    assert world.info().num_bricks == info["bricks_loaded"]
    ow = dense_world(oracle_mod, vox)
    for pos, yaw, pitch in [((20.5, 100.25, 30.75), 45.0, -30.0), ((110.3, 70.6, 120.2), 225.0, -15.0)]:
        cam = V.CameraSpec(pos, yaw, pitch, 640, 360, tile=(8, 8), anchor=anchor)
        rb = cam.rays_device()
        h = world.cast(rb, 128, with_pos=True)
        torch.cuda.synchronize()
        o, d = cam.rays_host()
        ref = ow.cast_simd(o, d, cam.world_origin, 128, with_pos=True)
        assert_hits_equal(h.t.cpu().numpy(), h.rec.cpu().numpy(), ref["t"], ref["rec"], f"{anchor} {pos}")
        assert np.array_equal(h.pos.cpu().numpy().view(np.uint32), ref["pos"].view(np.uint32))
        hit = (ref["rec"][:, 3] & 0xFF) != 0
        assert 0.3 < hit.mean() < 1.0
