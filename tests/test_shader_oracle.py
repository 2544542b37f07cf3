"""CPU pins of the oracle's shader method (VoxelTraversal.glsl rayCast / rayCastCoarse).

The shader method is not bit-compatible with the CPU cast, so it is pinned by what
the mathematics fixes: the ray-cell table by its closed form, exact (LOD-safe)
first hits by an FP64 brute-force DDA, the sector-of-sectors level by a hand-
counted iteration sequence, and the coarse variant by its invariants."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import blob_world
from test_oracle import dense_world


def camera_rays(rng, vox, n):
    """Rays from one empty voxel (a camera): W = floor(cam), origins fract(cam), unit directions."""
    sy, sz, sx = vox.shape
    while True:
        cam = rng.uniform(1, [sx - 1, sy - 1, sz - 1])
        if vox[int(cam[1]), int(cam[2]), int(cam[0])] == 0:
            break
    W = np.floor(cam).astype(np.int64)
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    d = v.astype(np.float32)
    d[d == 0] = np.float32(2.0**-24)
    o = np.repeat((cam - W).astype(np.float32)[:, None], n, axis=1)
    return np.ascontiguousarray(o), np.ascontiguousarray(d), tuple(int(x) for x in W)


def test_ray_cell_lut_closed_form(oracle_mod):
    """Entry (octant, cell c) holds exactly the cells q with (q_a - c_a) * s_a >= 0 on every axis."""
    lut = oracle_mod.ray_cell_lut()
    j = np.arange(64)
    q = np.stack([j & 3, (j >> 4) & 3, (j >> 2) & 3])  # x, y, z of each cell
    for oct in range(8):
        s = np.array([1 if oct & 1 else -1, 1 if oct & 2 else -1, 1 if oct & 4 else -1])[:, None]
        for c in range(64):
            cc = np.array([c & 3, (c >> 4) & 3, (c >> 2) & 3])[:, None]
            reach = np.all((q - cc) * s >= 0, axis=0)
            want = int(np.sum((reach.astype(np.uint64) << j.astype(np.uint64)), dtype=np.uint64))
            assert int(lut[oct * 64 + c]) == want, (oct, c)
    # every entry holds its own cell; the +++ octant from cell 0 reaches everything
    assert all((int(lut[o * 64 + c]) >> c) & 1 for o in range(8) for c in range(64))
    assert int(lut[7 * 64 + 0]) == (1 << 64) - 1


def test_sector_of_sectors_iteration_count(oracle_mod):
    """512^3 world holding one voxel at (500, 5, 5); a ray along +x from (0.5, 5.5, 5.5) steps
    x = 1 -> 127 (4x4x4 sector group empty: 128), 128 -> 255, 256 -> 383, 384 -> 447 (2x2x2
    sector group empty: 64), 448 -> 479 (32), 480 -> 495 (2x2x2 brick group empty: 16),
    496 -> 499 (4x4x4 voxel cell empty: 4), and hits at x = 500 with loop index 7."""
    vox = np.zeros((512, 512, 512), np.uint8)
    vox[5, 5, 500] = 9
    ow = dense_world(oracle_mod, vox)
    tiny = np.float32(2.0**-24)
    o = np.array([[0.5], [5.5], [5.5]], np.float32)
    d = np.array([[1.0], [tiny], [tiny]], np.float32)
    for lut in (False, True):
        r = ow.cast_shader(o, d, (0, 0, 0), "full", lut)
        x, y, z, info = (int(v) for v in r["rec"][0])
        assert (x, y, z) == (500, 5, 5)
        assert info & 0xFF == 9
        assert (info >> 16) == 7
        assert (info >> 8) & 63 == 3  # normal (-1, 0, 0)
        assert abs(float(r["t"][0]) - 499.5) < 1e-3


@pytest.mark.parametrize("lut", [False, True])
def test_full_method_first_hit_matches_fp64_dda(oracle_mod, lut):
    """rayCast is LOD-safe: its hit voxel is the first solid voxel after the start voxel, for
    every ray whose nearest pair of face crossings is separated by more than 1e-4."""
    rng = np.random.default_rng(11)
    vox = blob_world(256, seed=4, density=0.0005, blobs=10)
    vox[:128, :128, :128] = 0  # an empty 4x4x4 group of sectors
    ow = dense_world(oracle_mod, vox)
    checked = hits = 0
    iters = []
    for _ in range(4):
        o, d, W = camera_rays(rng, vox, 600)
        r = ow.cast_shader(o, d, W, "full", lut, with_pos=True)
        rec = r["rec"]
        assert np.array_equal(rec, ow.cast_shader(o, d, W, "full", lut)["rec"])
        hits_i = (rec[:, 3] & 0xFF) != 0
        # the reported position lies on the hit voxel (it is the point just past its entry face)
        p = r["pos"][:, hits_i].astype(np.float64) + np.array(W, np.float64)[:, None]
        assert np.all(np.abs(p - (rec[hits_i, :3].T + 0.5)) <= 0.5 + 1e-3)
        iters.append((rec[:, 3] >> 16) & 0xFFFF)
        for i in range(o.shape[1]):
            h, v, t, gap = ow.dda(o[:, i].astype(np.float64), d[:, i].astype(np.float64), W, 4096)
            if gap < 1e-4:
                continue
            checked += 1
            got = (rec[i, 3] & 0xFF) != 0
            assert got == h, f"ray {i}: shader hit={got}, dda hit={h}"
            if h:
                hits += 1
                assert tuple(int(x) for x in rec[i, :3]) == v
                assert abs(float(r["t"][i]) - t) <= 1e-4 * max(1.0, t)
    assert checked > 2000 and hits > 300
    assert np.concatenate(iters).max() < 256


def test_coarse_method_invariants(oracle_mod):
    """rayCastCoarse: hits before loop index 31 are exact first hits; later hits are solid voxels
    of a 4x4x4 cell that the ray passes through."""
    rng = np.random.default_rng(12)
    vox = blob_world(256, seed=5, density=0.002, blobs=8)
    ow = dense_world(oracle_mod, vox)
    late = early = 0
    for lut in (False, True):
        o, d, W = camera_rays(rng, vox, 800)
        r = ow.cast_shader(o, d, W, "coarse", lut)
        rec = r["rec"]
        for i in range(o.shape[1]):
            if (rec[i, 3] & 0xFF) == 0:
                continue
            it = (rec[i, 3] >> 16) & 0xFFFF
            v = tuple(int(x) for x in rec[i, :3])
            assert vox[v[1], v[2], v[0]] == (rec[i, 3] & 0xFF) != 0
            p = np.array(W, np.float64) + o[:, i].astype(np.float64)
            dd = d[:, i].astype(np.float64)
            if it <= 30:
                h, vv, t, gap = ow.dda(o[:, i].astype(np.float64), dd, W, 4096)
                if gap >= 1e-4:
                    early += 1
                    assert h and vv == v
            else:
                late += 1
                lo = (np.array(v) & ~3).astype(np.float64)
                t1, t2 = (lo - p) / dd, (lo + 4 - p) / dd
                assert np.max(np.minimum(t1, t2)) <= np.min(np.maximum(t1, t2)) + 1e-6
    assert early > 100 and late > 10
