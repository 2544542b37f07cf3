"""Pins of the oracle (the CPU restatement of CpuRenderer.cpp:110-224) against
closed forms, special cases, a brute-force FP64 DDA, the committed golden
fixture and its own SIMD twin.  CPU only."""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

from helpers import blob_world, random_rays

GOLDEN = Path(__file__).resolve().parent / "golden"


def dense_world(oracle_mod, vox: np.ndarray):
    """OracleWorld from a dense [Y][Z][X] grid (sector stream built here in numpy)."""
    sy, sz, sx = vox.shape
    lxz, ly = sx.bit_length() - 1, sy.bit_length() - 1
    nxz, ny = sx // 32, sy // 32
    idx, masks, bricks = [], [], []
    for cy in range(ny):
        for cz in range(nxz):
            for cx in range(nxz):
                sec = vox[cy * 32:(cy + 1) * 32, cz * 32:(cz + 1) * 32, cx * 32:(cx + 1) * 32]
                m = 0
                bl = []
                for b in range(64):
                    bx, bz, by = b & 3, (b >> 2) & 3, b >> 4
                    br = sec[by * 8:(by + 1) * 8, bz * 8:(bz + 1) * 8, bx * 8:(bx + 1) * 8]
                    if br.any():
                        m |= 1 << b
                        bl.append(br.reshape(-1))  # [y][z][x] flattening = x | z << 3 | y << 6
                if m:
                    idx.append(cx | cz << (nxz.bit_length() - 1) | cy << (2 * (nxz.bit_length() - 1)))
                    masks.append(m)
                    bricks.extend(bl)
    br = np.concatenate(bricks) if bricks else np.zeros(0, np.uint8)
    return oracle_mod.OracleWorld(lxz, ly, np.array(idx, np.uint32), np.array(masks, np.uint64), br)


def cast1(ow, o, d, W=(0, 0, 0), max_iters=128):
    r = ow.cast(np.array(o, np.float32).reshape(3, 1), np.array(d, np.float32).reshape(3, 1), W, max_iters,
                with_pos=True)
    rec = r["rec"][0]
    info = int(rec[3]) & 0xFFFFFFFF
    normal = tuple(((info >> (8 + 2 * a)) & 3 ^ 2) - 2 for a in range(3))
    return float(r["t"][0]), tuple(int(x) for x in rec[:3]), info & 0xFF, normal, info >> 16


# ---------------------------------------------------------------- world layout


def test_voxel_lookup_matches_dense(oracle_mod):
    vox = blob_world(64, seed=11)
    ow = dense_world(oracle_mod, vox)
    rng = np.random.default_rng(1)
    for p in rng.integers(0, 64, size=(5000, 3)):
        assert ow.voxel(*p) == vox[p[1], p[2], p[0]]
    assert ow.voxel(-1, 0, 0) == 0 and ow.voxel(0, 64, 0) == 0


# ---------------------------------------------------------------- special cases


def test_empty_world_counts_sector_cells(oracle_mod):
    """Empty world: every lookup sees mask 0 -> LOD 5 (32-voxel cells); the iteration
    count equals the number of 32^3 cells the ray passes through before leaving."""
    ow = dense_world(oracle_mod, np.zeros((128, 128, 128), np.uint8))
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(300):
        o = rng.uniform(1, 127, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o32, d32 = o.astype(np.float32), d.astype(np.float32)
        # FP64 closed form: cells crossed along each axis before the exit point
        od, dd = o32.astype(np.float64), d32.astype(np.float64)
        texit = min(((128.0 if dd[a] > 0 else 0.0) - od[a]) / dd[a] for a in range(3))
        pe = od + dd * texit
        # skip rays that pass within 1e-2 of a 32-boundary corner or exit near a cell edge
        near = False
        for a in range(3):
            for k in range(0, 129, 32):
                if abs(pe[a] - k) < 1e-2 and abs(((128.0 if dd[a] > 0 else 0.0) - k)) > 1e-9:
                    near = True
        if near:
            continue
        cells = 1 + sum(abs(math.floor(min(max(pe[a], 0.0), 127.999) / 32) - math.floor(od[a] / 32)) for a in range(3))
        t, v, mat, n, iters = cast1(ow, o32, d32)
        assert mat == 0 and t == -1.0
        assert iters == cells, (o, d, iters, cells)
        checked += 1
    assert checked > 200


def test_full_world_hits_at_origin(oracle_mod):
    """All voxels solid: hit at iteration 0, t = 0, voxel = W + floor(o), normal from the tie rule
    (sideDist all 0 -> X and Y set, Z clear; CpuRenderer.cpp:204-217)."""
    vox = np.full((64, 64, 64), 7, np.uint8)
    ow = dense_world(oracle_mod, vox)
    for o, d in [((10.5, 20.25, 30.75), (0.3, 0.5, 0.81)), ((1.5, 2.5, 3.5), (-0.6, 0.0, -0.8)),
                 ((63.9, 0.1, 5.0), (0.1, -0.9, 0.4))]:
        t, v, mat, n, iters = cast1(ow, o, d, W=(0, 0, 0))
        assert t == 0.0 and mat == 7 and iters == 1
        assert v == tuple(int(math.floor(x)) for x in o)
        exp = tuple((1 if math.copysign(1, c) < 0 else -1) for c in d[:2]) + (0,)
        assert n == exp


def test_axis_ray_single_voxel_closed_form(oracle_mod):
    """A lone voxel straight ahead: t is the distance to its entry face, exactly."""
    vox = np.zeros((64, 64, 64), np.uint8)
    vox[20, 20, 40] = 9  # (x, y, z) = (40, 20, 20)
    ow = dense_world(oracle_mod, vox)
    e = 2.0 ** -20
    t, v, mat, n, _ = cast1(ow, (3.25, 20.5, 20.5), (1.0, e, e))
    assert (t, v, mat, n) == (36.75, (40, 20, 20), 9, (-1, 0, 0))
    t, v, mat, n, _ = cast1(ow, (60.75, 20.5, 20.5), (-1.0, e, e))
    assert (t, v, mat, n) == (19.75, (40, 20, 20), 9, (1, 0, 0))
    # along y and z as well
    vox2 = np.zeros((64, 64, 64), np.uint8)
    vox2[50, 10, 12] = 3  # (12, 50, 10)
    ow2 = dense_world(oracle_mod, vox2)
    t, v, mat, n, _ = cast1(ow2, (12.5, 1.5, 10.5), (e, 1.0, e))
    assert (t, v, mat, n) == (48.5, (12, 50, 10), 3, (0, -1, 0))
    t, v, mat, n, _ = cast1(ow2, (12.5, 50.5, 63.0), (e, e, -1.0))
    assert (t, v, mat, n) == (52.0, (12, 50, 10), 3, (0, 0, 1))


def test_world_origin_shift_is_invariant(oracle_mod):
    """Rays are relative to W: shifting W and o by the same integer vector keeps absolute results."""
    vox = blob_world(64, seed=5)
    ow = dense_world(oracle_mod, vox)
    o, d = random_rays(2000, 64, seed=9)
    a = ow.cast(o, d, (0, 0, 0), 128)
    W = np.array([16, -8, 40], np.int32)
    b = ow.cast(o - W[:, None].astype(np.float32), d, tuple(W), 128)
    exact = np.all(np.abs(W) <= 64)
    assert exact
    # the float origins o - W are exact (small integers), but side distances are computed from
    # W-relative coordinates, so only voxel/material of non-grazing hits must agree
    same = np.all(a["rec"][:, :3] == b["rec"][:, :3], axis=1)
    assert same.mean() > 0.99


def test_exact_axis_aligned_ray_nan_path(oracle_mod):
    """d with exact zeros: 1/0 = inf, (rel * inf) + (-inf) = NaN side distances, min keeps the NaN
    (x86 minps returns the second operand), floor2i(NaN) = INT_MIN -> out of bounds: a miss after one lookup."""
    vox = np.zeros((64, 64, 64), np.uint8)
    vox[20, 20, 40] = 9
    ow = dense_world(oracle_mod, vox)
    t, v, mat, n, iters = cast1(ow, (3.25, 20.5, 20.5), (1.0, 0.0, 0.0))
    assert mat == 0 and iters == 1


def test_origin_outside_and_iteration_cap(oracle_mod):
    vox = blob_world(64, seed=2)
    ow = dense_world(oracle_mod, vox)
    t, v, mat, n, iters = cast1(ow, (-5.5, 10.0, 10.0), (1.0, 0.01, 0.01))
    assert mat == 0 and iters == 0
    t, v, mat, n, iters = cast1(ow, (10.5, 63.5, 10.5), (0.001, -1.0, 0.002), max_iters=3)
    full = cast1(ow, (10.5, 63.5, 10.5), (0.001, -1.0, 0.002), max_iters=128)
    if full[4] > 3:
        assert mat == 0 and iters == 3


# ---------------------------------------------------------------- brute force


def test_first_solid_voxel_matches_fp64_dda(oracle_mod):
    """On random sparse worlds the LOD traversal returns the first solid voxel a plain FP64
    voxel walk finds (Amanatides-Woo, no LOD), for every ray whose path does not graze an edge
    within 3e-3 (the +0.001 bias may skip such slivers) and that did not exhaust the cap."""
    for seed in (1, 2, 3):
        vox = blob_world(64, seed=seed, density=0.004, blobs=4)
        ow = dense_world(oracle_mod, vox)
        o, d = random_rays(4000, 64, seed=100 + seed)
        r = ow.cast(o, d, (0, 0, 0), 512)
        compared = 0
        for i in range(o.shape[1]):
            hit, v, t, gap = ow.dda(o[:, i].astype(np.float64), d[:, i].astype(np.float64), (0, 0, 0), 4096)
            if gap < 3e-3:
                continue
            rec = r["rec"][i]
            iters = (int(rec[3]) >> 16) & 0xFFFF
            if iters >= 512:
                continue
            ohit = (rec[3] & 0xFF) != 0
            assert ohit == hit, (seed, i)
            if hit:
                assert tuple(rec[:3]) == v, (seed, i, rec, v)
                assert abs(float(r["t"][i]) - t) <= 1e-4 * max(1.0, t)
                assert (rec[3] & 0xFF) == vox[v[1], v[2], v[0]]
            compared += 1
        assert compared > 3000


def test_simd_twin_is_bit_exact(oracle_mod):
    vox = blob_world(128, seed=4, density=0.01)
    ow = dense_world(oracle_mod, vox)
    o, d = random_rays(20000, 128, seed=8, inside=False)
    a = ow.cast(o, d, (0, 0, 0), 128)
    b = ow.cast_simd(o, d, (0, 0, 0), 128, threads=2)
    assert np.array_equal(a["rec"], b["rec"])
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))


def test_fp64_picking_cast_matches_dda(oracle_mod):
    """VoxelMap::RayCast (VoxelMap.cpp:140-170) restated: it steps out of the origin voxel first,
    then returns the first solid voxel with distance = entry + 1e-4 bias."""
    vox = blob_world(64, seed=7, density=0.004, blobs=3)
    ow = dense_world(oracle_mod, vox)
    o, d = random_rays(1500, 64, seed=77)
    compared = 0
    for i in range(o.shape[1]):
        oo, dd = o[:, i].astype(np.float64), d[:, i].astype(np.float64)
        if vox[int(oo[1]), int(oo[2]), int(oo[0])]:
            continue  # the picking cast never tests the origin voxel
        hit, v, t, gap = ow.dda(oo, dd, (0, 0, 0), 4096)
        if gap < 1e-3:
            continue
        r = ow.cast_f64(oo, dd, 1024)
        if hit:
            assert r["voxel"] == v and abs(r["distance"] - (t + 1e-4)) < 1e-9
            nz = [a for a in range(3) if r["normal"][a] != 0]
            assert len(nz) == 1
        else:
            # out of the world: the picking cast keeps walking empty sectors until maxIters
            assert r["distance"] < 0
        compared += 1
    assert compared > 1000


# ---------------------------------------------------------------- bounce inputs


def test_sincos_turns_accuracy(oracle_mod):
    u = np.linspace(0.0, 1.0, 20001, dtype=np.float32)
    s = np.array([oracle_mod.sincos_turns(float(x)) for x in u])
    ang = 2 * np.pi * u.astype(np.float64)
    assert np.max(np.abs(s[:, 0] - np.sin(ang))) < 3e-7
    assert np.max(np.abs(s[:, 1] - np.cos(ang))) < 3e-7


def test_pcg_stream_is_uniform_and_seeded(oracle_mod):
    vals = []
    for px in range(64):
        for py in range(64):
            st = oracle_mod.pcg_init(px, py, 1)
            v, st = oracle_mod.pcg_next(st)
            vals.append(v)
    vals = np.array(vals, np.float64) / 2**32
    hist, _ = np.histogram(vals, bins=16, range=(0, 1))
    chi2 = ((hist - len(vals) / 16) ** 2 / (len(vals) / 16)).sum()
    assert chi2 < 45  # 15 dof, p ~ 1e-4
    assert oracle_mod.pcg_init(3, 4, 1) != oracle_mod.pcg_init(3, 4, 2)
    # the state advances by the LCG of RandomGen.glsl:9
    s0 = oracle_mod.pcg_init(5, 6, 7)
    _, s1 = oracle_mod.pcg_next(s0)
    assert s1 == (s0 * 747796405 + 2891336453) % 2**32


def test_bounce_is_cosine_distributed(oracle_mod):
    n = 40000
    rng = np.random.default_rng(5)
    normals = [(1, 0, 0), (3, 0, 0), (0, 1, 0), (0, 3, 0), (0, 0, 1), (0, 0, 3)]  # codes +1 / -1
    code = np.array([normals[k] for k in rng.integers(0, 6, n)])
    info = (7 | (code[:, 0] << 8) | (code[:, 1] << 10) | (code[:, 2] << 12)).astype(np.int64)
    rec = np.zeros((n, 4), np.int32)
    rec[:, 3] = info.astype(np.int32)
    pos = rng.uniform(0, 64, (3, n)).astype(np.float32)
    d = np.zeros((3, n), np.float32)
    oo, od, src = oracle_mod.bounce(rec, pos, d, 256, 1)
    assert len(src) == n
    N = np.where(code == 3, -1, code).T.astype(np.float64)
    cosv = np.sum(N * od, axis=0)
    assert np.all(cosv >= -1e-6)
    assert abs(cosv.mean() - 2 / 3) < 0.01  # E[N.d] of a cosine-weighted hemisphere
    assert np.max(np.abs(np.linalg.norm(od, axis=0) - 1)) < 1e-6
    assert np.array_equal(oo, (pos + (N * 0.01).astype(np.float32)).astype(np.float32))


# ---------------------------------------------------------------- golden fixture


def test_golden_fixture(oracle_mod):
    """The committed fixture (written by tests/golden/make_golden.py) is reproduced bit-exactly."""
    f = np.load(GOLDEN / "blob64.npz")
    ow = dense_world(oracle_mod, f["voxels"])
    r = ow.cast(f["o"], f["d"], tuple(f["world_origin"]), int(f["max_iters"]))
    assert np.array_equal(r["rec"], f["rec"])
    assert np.array_equal(r["t"].view(np.uint32), f["t"].view(np.uint32))


def test_uv_closed_form(oracle_mod):
    """UV of a hit (CpuRenderer.cpp:217-220): on an x face (u, v) = fract(y, z), on a y face
    fract(x, z), on a z face fract(x, y); at an x/y edge tie the x rule.  Nearly axis-aligned
    rays hit a lone voxel's faces at known coordinates."""
    vox = np.zeros((64, 64, 64), np.uint8)
    vox[20, 20, 40] = 9  # (x, y, z) = (40, 20, 20)
    ow = dense_world(oracle_mod, vox)
    e = 2.0 ** -20
    cases = [  # origin, direction, expected (u, v)
        ((3.0, 20.25, 20.625), (1.0, e, e), (0.25, 0.625)),        # x face: (y, z)
        ((40.375, 2.0, 20.75), (e, 1.0, e), (0.375, 0.75)),        # y face: (x, z)
        ((40.125, 20.875, 60.0), (e, e, -1.0), (0.125, 0.875)),    # z face: (x, y)
    ]
    for o, d, (u, v) in cases:
        r = ow.cast(np.array(o, np.float32).reshape(3, 1), np.array(d, np.float32).reshape(3, 1), (0, 0, 0), 128,
                    with_pos=True, with_uv=True)
        assert (r["rec"][0, 3] & 0xFF) == 9
        assert abs(float(r["uv"][0, 0]) - u) < 1e-4 and abs(float(r["uv"][1, 0]) - v) < 1e-4, (o, r["uv"][:, 0])
    # a miss has uv 0
    r = ow.cast(np.array([[3.0], [50.5], [3.5]], np.float32), np.array([[1.0], [e], [e]], np.float32), (0, 0, 0),
                128, with_uv=True)
    assert (r["rec"][0, 3] & 0xFF) == 0 and r["uv"][0, 0] == 0 and r["uv"][1, 0] == 0
    # all voxels solid: the ray starts in one, side distances tie at 0 on every axis: X and Y set,
    # so u = fract(y), v = fract(z)
    ow = dense_world(oracle_mod, np.full((64, 64, 64), 5, np.uint8))
    r = ow.cast(np.array([[10.5], [20.25], [30.75]], np.float32), np.array([[0.3], [0.5], [0.81]], np.float32),
                (0, 0, 0), 128, with_pos=True, with_uv=True)
    assert float(r["uv"][0, 0]) == np.float32(0.25) and float(r["uv"][1, 0]) == np.float32(0.75)
