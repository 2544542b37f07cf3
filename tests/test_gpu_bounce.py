"""GPU parity of the primary + diffuse pipeline (vrt_cast -> vrt_bounce_generate -> vrt_cast)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from helpers import assert_hits_equal

pytestmark = pytest.mark.gpu


def run_pipeline(V, world, prim, width, seed, max_iters, schedule=1):
    hits = world.cast(prim, max_iters, with_pos=True, schedule=1)
    br, src, count = V.bounce(prim, hits, width, seed)
    hb = world.cast(br, max_iters, schedule=schedule, count=count)
    torch.cuda.synchronize()
    nb = int(count.item())
    return hits, br, src, nb, hb


@pytest.fixture(scope="module")
def c2_scene():
    from voxelrt_amd import workloads as W

    return W.C2_WORLD.build()


# C2's camera rays with their integer world origin at the camera's sector corner (the bench's
# aligned-origin step), at floor(eye) (the reference's anchor, CpuRenderer.cpp:447; the general
# step) and at 0 (world-space rays; the zero-origin step)
@pytest.mark.parametrize("config", ["c1", "c2", "c2_floor", "c2_zero"])
def test_primary_and_bounce_parity(vrt, oracle_mod, cuda, config, c2_scene):
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    if config == "c1":
        scene, cam = W.C1_WORLD.build(), W.c1_camera()
    else:
        anchor = {"c2": "sector", "c2_floor": "floor", "c2_zero": "zero"}[config]
        scene, cam = c2_scene, W.c2_camera(anchor=anchor)
    world = V.World.from_scene(scene)
    ow = oracle_mod.OracleWorld.from_stream(scene.stream())
    prim = cam.rays_device()
    hits, br, src, nb, hb = run_pipeline(V, world, prim, cam.c.width, W.BOUNCE_SEED, W.MAX_ITERS,
                                         schedule=2 if config == "c1" else 1)
    print(f"{config}: world origin {prim.world_origin}, {prim.n} primary and {nb} bounce rays", flush=True)
    po, pd = prim.o.cpu().numpy(), prim.d.cpu().numpy()
    ref = ow.cast_simd(po, pd, prim.world_origin, W.MAX_ITERS, with_pos=True)
    g_t, g_rec, g_pos = hits.t.cpu().numpy(), hits.rec.cpu().numpy(), hits.pos.cpu().numpy()
    assert_hits_equal(g_t, g_rec, ref["t"], ref["rec"], f"{config} primary")
    # every primary's hit position (0 on a miss), bit for bit
    assert np.array_equal(g_pos.view(np.uint32), ref["pos"].view(np.uint32))
    # the other schedules and block order on the full set: the persistent queue (two chunk
    # sizes) and the grid dispatched last block first
    for sched, chunk, order in ((2, 256, 0), (2, 1024, 0), (1, 0, 1)):
        h2 = world.cast(prim, W.MAX_ITERS, with_pos=True, schedule=sched, chunk=chunk, order=order)
        torch.cuda.synchronize()
        what = f"{config} primary schedule={sched} chunk={chunk} order={order}"
        assert_hits_equal(h2.t.cpu().numpy(), h2.rec.cpu().numpy(), ref["t"], ref["rec"], what)
        assert np.array_equal(h2.pos.cpu().numpy().view(np.uint32), ref["pos"].view(np.uint32)), what
    # the SIMD twin's positions equal the scalar oracle's on a slice of the rays
    ref_s = ow.cast(po, pd, prim.world_origin, W.MAX_ITERS, with_pos=True, begin=0, end=min(po.shape[1], 50000))
    assert np.array_equal(ref_s["pos"].view(np.uint32), ref["pos"][:, :ref_s["pos"].shape[1]].view(np.uint32))

    # bounce rays from the oracle's own hits: oracle restatement, the library's host twin and the
    # device kernel agree bit-exactly
    oo, od, osrc = oracle_mod.bounce(ref["rec"], ref["pos"], pd, cam.c.width, W.BOUNCE_SEED)
    assert nb == len(osrc)
    assert np.array_equal(src[:nb].cpu().numpy().astype(np.uint32), osrc)
    assert np.array_equal(br.o[:, :nb].cpu().numpy().view(np.uint32), oo.view(np.uint32))
    assert np.array_equal(br.d[:, :nb].cpu().numpy().view(np.uint32), od.view(np.uint32))

    # bounce casts
    rb = ow.cast_simd(oo, od, prim.world_origin, W.MAX_ITERS)
    assert_hits_equal(hb.t[:nb].cpu().numpy(), hb.rec[:nb].cpu().numpy(), rb["t"], rb["rec"], f"{config} bounce")
    # the compact schedule on the bounce set (device count, capacity-sized first pass)
    count = torch.tensor([nb], dtype=torch.int32, device="cuda")
    for cap in (0, 24):
        hc = world.cast(br, W.MAX_ITERS, schedule=3, pass_cap=cap, count=count)
        torch.cuda.synchronize()
        assert_hits_equal(hc.t[:nb].cpu().numpy(), hc.rec[:nb].cpu().numpy(), rb["t"], rb["rec"],
                          f"{config} bounce compact cap={cap}")
        assert np.array_equal(hc.t[:nb].cpu().numpy().view(np.uint32), rb["t"].view(np.uint32))
    # bounce rays start next to a solid face and go away from it: cosine-weighted around N
    n_code = g_rec[osrc, 3]
    N = np.stack([(((n_code >> (8 + 2 * a)) & 3) ^ 2) - 2 for a in range(3)]).astype(np.float64)
    N /= np.linalg.norm(N, axis=0)
    cosv = np.sum(N * od, axis=0)
    assert cosv.min() > -1e-6
    assert abs(cosv.mean() - 2 / 3) < 0.02


def test_host_twin_of_bounce(vrt, cuda):
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L
    from voxelrt_amd import workloads as W

    scene, cam = W.C1_WORLD.build(), W.c1_camera()
    world = V.World.from_scene(scene)
    prim = cam.rays_device()
    hits, br, src, nb, hb = run_pipeline(V, world, prim, cam.c.width, 7, W.MAX_ITERS)
    n = prim.n
    rec = np.ascontiguousarray(hits.rec.cpu().numpy())
    pos = np.ascontiguousarray(hits.pos.cpu().numpy())
    out = np.zeros((6, n), np.float32)
    hsrc = np.zeros(n, np.uint32)
    cnt = C.c_uint32()
    ptrs = (C.c_void_p * 6)(*[out[i].ctypes.data for i in range(6)])
    L.check(V.lib().vrt_bounce_host(n, rec.ctypes.data, pos.ctypes.data, cam.c.width, 7, C.byref(ptrs),
                                    hsrc.ctypes.data, C.byref(cnt)))
    assert cnt.value == nb
    assert np.array_equal(out[:3, :nb].view(np.uint32), br.o[:, :nb].cpu().numpy().view(np.uint32))
    assert np.array_equal(out[3:, :nb].view(np.uint32), br.d[:, :nb].cpu().numpy().view(np.uint32))


def test_device_count_bounds_the_cast(vrt, oracle_mod, cuda):
    """A cast given a device count traces exactly that many rays and leaves the rest untouched."""
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    scene, cam = W.C1_WORLD.build(), W.c1_camera()
    world = V.World.from_scene(scene)
    prim = cam.rays_device()
    for sched in (1, 2, 3):
        for k in (0, 1, 777, prim.n):
            cnt = torch.tensor([k], dtype=torch.int32, device="cuda")
            h = V.HitBatch.empty(prim.n)
            h.rec.fill_(-7)
            ctr = V.new_counters()
            world.cast(prim, W.MAX_ITERS, schedule=sched, count=cnt, out=h, counters=ctr)
            torch.cuda.synchronize()
            rec = h.rec.cpu().numpy()
            assert int(ctr[0].item()) == k
            assert np.all(rec[k:] == -7)
            assert not np.any(np.all(rec[:k] == -7, axis=1))


def test_bounce_compaction_of_many_blocks(vrt, cuda):
    """More than 15872 blocks of 256 primaries (the scan's LDS staging limit): the scan takes
    its global-memory path; the compacted bounce set equals the host twin."""
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L
    from voxelrt_amd import workloads as W

    scene = W.C1_WORLD.build()
    world = V.World.from_scene(scene)
    cam = W.c1_camera(tile=(8, 8), width=2560, height=1600)   # 4 096 000 rays, 16 000 blocks
    prim = cam.rays_device()
    hits, br, src, nb, hb = run_pipeline(V, world, prim, cam.c.width, 5, W.MAX_ITERS)
    n = prim.n
    rec = np.ascontiguousarray(hits.rec.cpu().numpy())
    pos = np.ascontiguousarray(hits.pos.cpu().numpy())
    out = np.zeros((6, n), np.float32)
    hsrc = np.zeros(n, np.uint32)
    cnt = C.c_uint32()
    ptrs = (C.c_void_p * 6)(*[out[i].ctypes.data for i in range(6)])
    L.check(V.lib().vrt_bounce_host(n, rec.ctypes.data, pos.ctypes.data, cam.c.width, 5, C.byref(ptrs),
                                    hsrc.ctypes.data, C.byref(cnt)))
    assert cnt.value == nb and nb > 1_000_000
    assert np.array_equal(src[:nb].cpu().numpy().astype(np.uint32), hsrc[:nb])
    assert np.array_equal(out[3:, :nb].view(np.uint32), br.d[:, :nb].cpu().numpy().view(np.uint32))


@pytest.mark.parametrize("n,frac", [(1, 1.0), (4095, 0.5), (4097, 0.3), (1_000_003, 0.45), (8_388_608, 0.5),
                                    (8_388_609, 0.02), (9_000_000, 0.6)])
def test_bounce_generation_paths_match_host_twin(vrt, cuda, n, frac):
    """Synthetic hit sets on both sides of the two-launch generation's limit (2048 blocks of
    4096 primaries; above it count -> scan -> scatter): rays, sources and count equal the
    host twin's, with ragged last blocks and sparse and dense hits."""
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L

    rng = np.random.default_rng(n)
    rec = np.zeros((n, 4), np.int32)
    hit = rng.random(n) < frac
    codes = rng.integers(0, 6, n)  # one axis, either sign: the 2-bit normal fields at bits 8/10/12
    info = np.where(hit, 1 + rng.integers(0, 255, n), 0)
    field = (np.where(codes % 2 == 0, 1, 3) << (8 + 2 * (codes // 2)))
    rec[:, 3] = np.where(hit, info | field, 0).astype(np.int32)
    pos = (rng.random((3, n), dtype=np.float32) * 64.0).astype(np.float32)
    hits = V.HitBatch(torch.zeros(n, device=cuda), torch.from_numpy(rec).to(cuda), torch.from_numpy(pos).to(cuda))
    prim = V.RayBatch.empty(n, (3, -5, 7))
    br, src, count = V.bounce(prim, hits, 1920, 7)
    torch.cuda.synchronize()
    nb = int(count.item())
    out = np.zeros((6, n), np.float32)
    hsrc = np.zeros(n, np.uint32)
    cnt = C.c_uint32()
    ptrs = (C.c_void_p * 6)(*[out[i].ctypes.data for i in range(6)])
    L.check(V.lib().vrt_bounce_host(n, rec.ctypes.data, pos.ctypes.data, 1920, 7, C.byref(ptrs), hsrc.ctypes.data,
                                    C.byref(cnt)))
    assert cnt.value == nb == int(hit.sum())
    assert np.array_equal(src[:nb].cpu().numpy().astype(np.uint32), hsrc[:nb])
    assert np.array_equal(out[:3, :nb].view(np.uint32), br.o[:, :nb].cpu().numpy().view(np.uint32))
    assert np.array_equal(out[3:, :nb].view(np.uint32), br.d[:, :nb].cpu().numpy().view(np.uint32))
    if n <= 8388608:
        # direction order: the same rays, grouped by key, primary order within a key
        b1, s1, c1 = V.bounce(prim, hits, 1920, 7, order=1)
        torch.cuda.synchronize()
        assert int(c1.item()) == nb
        src1 = s1[:nb].cpu().numpy().astype(np.int64)
        j = np.searchsorted(hsrc[:nb].astype(np.int64), src1)
        assert np.array_equal(hsrc[:nb][j].astype(np.int64), src1)
        assert np.array_equal(b1.d[:, :nb].cpu().numpy().view(np.uint32), out[3:, :nb][:, j].view(np.uint32))
        assert np.array_equal(b1.o[:, :nb].cpu().numpy().view(np.uint32), out[:3, :nb][:, j].view(np.uint32))
        key = _bounce_keys(rec[src1, 3], src1, 1920, 7)
        assert np.all(np.diff(key) >= 0) and np.all(np.diff(src1)[np.diff(key) == 0] > 0)


def _bounce_keys(info, src, width, seed):
    """Host model of the direction-order key: the PCG samples (RandomGen.glsl:6-20) of primary src
    and the signs of N + s's terms (s = (sin 2 pi u2 sy, 2 u1 - 1, cos 2 pi u2 sy))."""
    M = np.uint32
    px, py = (src % width).astype(M), (src // width).astype(M)
    with np.errstate(over="ignore"):
        st = ((M(1103515245) * px) ^ ((M(1103515245) * py) >> M(3))) + M(seed) * M(12345)

        def nxt(s):
            word = ((s >> ((s >> M(28)) + M(4))) ^ s) * M(277803737)
            return s * M(747796405) + M(2891336453), (word >> M(22)) ^ word

        st, w1 = nxt(st)
        st, w2 = nxt(st)
    u1 = w1.astype(np.float32) * np.float32(2.0 ** -32)
    u2 = w2.astype(np.float32) * np.float32(2.0 ** -32)
    info = info.astype(np.uint32)
    cx, cy, cz = (info >> 8) & 3, (info >> 10) & 3, (info >> 12) & 3
    bx = np.where(cx != 0, cx == 1, u2 < 0.5)
    by = np.where(cy != 0, cy == 1, u1 >= 0.5)
    bz = np.where(cz != 0, cz == 1, (u2 < 0.25) | (u2 >= 0.75))
    return bx.astype(np.int64) | by.astype(np.int64) << 1 | bz.astype(np.int64) << 2


@pytest.mark.parametrize("config", ["c1", "c2"])
def test_direction_order_generation(vrt, cuda, config, c2_scene):
    """vrt_bounce_generate_ex in direction order writes the primary-order set's rays, sources and
    count, permuted: grouped by key (the host model of the key), increasing primary index within
    a key; the key is the octant of the written direction for all but sign-boundary rays; the
    cast of either order gives each ray the same hit."""
    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    scene, cam = (W.C1_WORLD.build(), W.c1_camera()) if config == "c1" else (c2_scene, W.c2_camera())
    world = V.World.from_scene(scene)
    prim = cam.rays_device()
    hits = world.cast(prim, W.MAX_ITERS, with_pos=True)
    b0, s0, c0 = V.bounce(prim, hits, cam.c.width, W.BOUNCE_SEED)
    b1, s1, c1 = V.bounce(prim, hits, cam.c.width, W.BOUNCE_SEED, order=1)
    torch.cuda.synchronize()
    nb = int(c0.item())
    assert int(c1.item()) == nb and nb > 0
    src0, src1 = s0[:nb].cpu().numpy().astype(np.int64), s1[:nb].cpu().numpy().astype(np.int64)
    assert np.all(np.diff(src0) > 0)
    j = np.searchsorted(src0, src1)
    assert np.array_equal(src0[j], src1) and len(np.unique(src1)) == nb
    for a in ("o", "d"):
        g0, g1 = getattr(b0, a)[:, :nb].cpu().numpy(), getattr(b1, a)[:, :nb].cpu().numpy()
        assert np.array_equal(g1.view(np.uint32), g0[:, j].view(np.uint32)), a
    info = hits.rec[:, 3].cpu().numpy()[src1]
    key = _bounce_keys(info, src1, cam.c.width, W.BOUNCE_SEED)
    assert np.all(np.diff(key) >= 0)
    same_key = np.diff(key) == 0
    assert np.all(np.diff(src1)[same_key] > 0)
    d1 = b1.d[:, :nb].cpu().numpy()
    octant = (d1[0] >= 0).astype(np.int64) | (d1[1] >= 0).astype(np.int64) << 1 | (d1[2] >= 0).astype(np.int64) << 2
    assert (octant == key).mean() > 0.9999
    h0 = world.cast(b0, W.MAX_ITERS, count=c0)
    h1 = world.cast(b1, W.MAX_ITERS, count=c1)
    torch.cuda.synchronize()
    r0, r1 = h0.rec[:nb].cpu().numpy(), h1.rec[:nb].cpu().numpy()
    assert np.array_equal(r1, r0[j])
    assert np.array_equal(h1.t[:nb].cpu().numpy().view(np.uint32), h0.t[:nb].cpu().numpy()[j].view(np.uint32))


def test_direction_order_limits(vrt, cuda):
    """Direction order takes at most 8388608 primaries (rejected before any launch), and no
    primaries writes a zero count."""
    import voxelrt_amd as V

    empty = V.RayBatch.empty(0)
    br, src, cnt = V.bounce(empty, V.HitBatch.empty(0, with_pos=True), 16, 1, order=1)
    torch.cuda.synchronize()
    assert int(cnt.item()) == 0
    with pytest.raises(V.VrtError, match="unknown order"):
        V.bounce(empty, V.HitBatch.empty(0, with_pos=True), 16, 1, order=2)
    n = 8388609
    with pytest.raises(V.VrtError, match="at most 8388608"):
        V.bounce(V.RayBatch.empty(n), V.HitBatch.empty(n, with_pos=True), 4096, 1, order=1)


@pytest.mark.parametrize("width", [1, 3, 640, 4097, 65537, 2**31 + 1, 2**32 - 1])
def test_bounce_generation_any_width(vrt, cuda, width):
    """The device splits the primary index into (i % width, i / width) with a multiply-high by a
    per-launch constant: rays and sources equal the host twin's (plain division) at widths
    from 1 to 2^32 - 1, in both orders."""
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L

    n = 20011
    rng = np.random.default_rng(width % 1000)
    rec = np.zeros((n, 4), np.int32)
    hit = rng.random(n) < 0.6
    codes = rng.integers(0, 6, n)
    info = np.where(hit, 1 + rng.integers(0, 255, n), 0)
    rec[:, 3] = np.where(hit, info | (np.where(codes % 2 == 0, 1, 3) << (8 + 2 * (codes // 2))), 0).astype(np.int32)
    pos = (rng.random((3, n), dtype=np.float32) * 64.0).astype(np.float32)
    hits = V.HitBatch(torch.zeros(n, device=cuda), torch.from_numpy(rec).to(cuda), torch.from_numpy(pos).to(cuda))
    prim = V.RayBatch.empty(n, (0, 0, 0))
    out = np.zeros((6, n), np.float32)
    hsrc = np.zeros(n, np.uint32)
    cnt = C.c_uint32()
    ptrs = (C.c_void_p * 6)(*[out[i].ctypes.data for i in range(6)])
    L.check(V.lib().vrt_bounce_host(n, rec.ctypes.data, pos.ctypes.data, width, 9, C.byref(ptrs), hsrc.ctypes.data,
                                    C.byref(cnt)))
    nb = cnt.value
    for order in (0, 1):
        br, src, count = V.bounce(prim, hits, width, 9, order=order)
        torch.cuda.synchronize()
        assert int(count.item()) == nb
        s = src[:nb].cpu().numpy().astype(np.int64)
        j = np.searchsorted(hsrc[:nb].astype(np.int64), s)
        assert np.array_equal(hsrc[:nb][j].astype(np.int64), s)
        assert np.array_equal(br.o[:, :nb].cpu().numpy().view(np.uint32), out[:3, :nb][:, j].view(np.uint32))
        assert np.array_equal(br.d[:, :nb].cpu().numpy().view(np.uint32), out[3:, :nb][:, j].view(np.uint32))
