"""GPU checks of the device world: incremental sector sync (the DirtyLocs protocol of
GpuRenderer.cpp:45-167), slot exhaustion, error codes and forced HIP failures."""
from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from helpers import assert_hits_equal, blob_world, random_rays

pytestmark = pytest.mark.gpu


def sector_stream_of(vox: np.ndarray, sectors: list[int]):
    """Sector stream (idx, mask, bricks) of the listed sector indices of a dense [Y][Z][X] grid."""
    sy, sz, sx = vox.shape
    nxz = sx // 32
    sh = nxz.bit_length() - 1
    idx, masks, bricks = [], [], []
    for si in sectors:
        cx, cz, cy = si & (nxz - 1), (si >> sh) & (nxz - 1), si >> (2 * sh)
        sec = vox[cy * 32:(cy + 1) * 32, cz * 32:(cz + 1) * 32, cx * 32:(cx + 1) * 32]
        m = 0
        for b in range(64):
            bx, bz, by = b & 3, (b >> 2) & 3, b >> 4
            br = sec[by * 8:(by + 1) * 8, bz * 8:(bz + 1) * 8, bx * 8:(bx + 1) * 8]
            if br.any():
                m |= 1 << b
                bricks.append(br.reshape(-1))
        idx.append(si)
        masks.append(m)
    br = np.concatenate(bricks) if bricks else np.zeros(0, np.uint8)
    return np.array(idx, np.uint32), np.array(masks, np.uint64), br


def cast_np(world, o, d, W=(0, 0, 0), max_iters=128):
    from voxelrt_amd import RayBatch

    h = world.cast(RayBatch.from_numpy(o, d, W), max_iters)
    torch.cuda.synchronize()
    return h.t.cpu().numpy(), h.rec.cpu().numpy()


class _DeviceWords:
    """A device buffer of u32 words, for torch.as_tensor."""

    def __init__(self, ptr: int, shape: tuple):
        self.__cuda_array_interface__ = {"shape": shape, "typestr": "<u4", "data": (ptr, False), "version": 3}


def sector_records(world) -> np.ndarray:
    """The world's device sector records as [num_sectors, 4] u32: mask lo, mask hi, base, group cell mask."""
    n = world.info().num_sectors
    torch.cuda.synchronize()
    return torch.as_tensor(_DeviceWords(world.buffers().sectors, (n, 4)), device="cuda").cpu().numpy()


def check_sector_levels(world, masks_by_sector: dict):
    rec = sector_records(world)
    mask = rec[:, 0].astype(np.uint64) | rec[:, 1].astype(np.uint64) << np.uint64(32)
    for si, m in masks_by_sector.items():
        assert int(mask[si]) == int(m), si
    # cell mask of a step over an empty brick group: 16 voxels, the whole sector when empty
    assert np.array_equal(rec[:, 3], np.where(mask == 0, 31, 15).astype(np.uint32))


def test_incremental_sync_equals_fresh_world(vrt, oracle_mod, cuda):
    import voxelrt_amd as V
    from test_oracle import dense_world

    rng = np.random.default_rng(3)
    vox = blob_world(128, seed=21, density=0.01)
    world = V.World(7, 7)
    nsec = 4 ** 3
    check_sector_levels(world, {si: 0 for si in range(nsec)})
    world.sync(*sector_stream_of(vox, list(range(nsec))))
    check_sector_levels(world, dict(zip(range(nsec), sector_stream_of(vox, list(range(nsec)))[1])))
    o, d = random_rays(20000, 128, seed=5)
    for step in range(4):
        # edit: clear a box, paint another, empty a whole sector
        v2 = vox.copy()
        a = rng.integers(0, 100, 3)
        v2[a[1]:a[1] + 20, a[2]:a[2] + 20, a[0]:a[0] + 20] = 0
        b = rng.integers(0, 110, 3)
        v2[b[1]:b[1] + 12, b[2]:b[2] + 12, b[0]:b[0] + 12] = rng.integers(1, 256)
        c = rng.integers(0, 4, 3)
        v2[c[1] * 32:(c[1] + 1) * 32, c[2] * 32:(c[2] + 1) * 32, c[0] * 32:(c[0] + 1) * 32] = 0
        changed = []
        for si in range(nsec):
            cx, cz, cy = si & 3, (si >> 2) & 3, si >> 4
            sl = np.s_[cy * 32:(cy + 1) * 32, cz * 32:(cz + 1) * 32, cx * 32:(cx + 1) * 32]
            if not np.array_equal(vox[sl], v2[sl]):
                changed.append(si)
        rng.shuffle(changed)
        world.sync(*sector_stream_of(v2, changed))
        vox = v2
        ref = dense_world(oracle_mod, vox).cast(o, d, (0, 0, 0), 128)
        t, rec = cast_np(world, o, d)
        assert_hits_equal(t, rec, ref["t"], ref["rec"], f"edit {step}")
        fresh = V.World.from_scene(V.Scene.from_dense(vox))
        t2, rec2 = cast_np(fresh, o, d)
        assert np.array_equal(rec, rec2)
        masks = sector_stream_of(vox, list(range(nsec)))[1]
        assert (masks == 0).any()  # the emptied sector
        check_sector_levels(world, dict(zip(range(nsec), masks)))
        info = world.info()
        assert info.num_bricks == sum(bin(int(m)).count("1") for m in sector_stream_of(vox, list(range(nsec)))[1])


def test_slot_exhaustion_raises_and_keeps_world(vrt, cuda):
    import voxelrt_amd as V

    vox = blob_world(64, seed=8, density=0.0, blobs=3)
    idx, masks, bricks = sector_stream_of(vox, list(range(8)))
    need = sum(bin(int(m)).count("1") for m in masks)
    assert 0 < need < 512
    world = V.World(6, 6, max_slots=need)
    world.sync(idx, masks, bricks)
    o, d = random_rays(4000, 64, seed=2)
    before = cast_np(world, o, d)
    full = np.full((64, 64, 64), 3, np.uint8)
    with pytest.raises(V.CapacityError, match="Could not allocate brick slots"):
        world.sync(*sector_stream_of(full, list(range(8))))
    after = cast_np(world, o, d)
    assert np.array_equal(before[1], after[1])
    assert world.info().num_bricks == need


def test_error_codes(vrt, cuda):
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L

    world = V.World(6, 6)
    with pytest.raises(V.VrtError) as e:
        world.sync(np.array([1, 1], np.uint32), np.array([1, 1], np.uint64), np.zeros(1024, np.uint8))
    assert e.value.status == L.VRT_ERR_INVALID_ARGUMENT and "twice" in str(e.value)
    with pytest.raises(V.VrtError) as e:
        world.sync(np.array([9], np.uint32), np.array([1], np.uint64), np.zeros(512, np.uint8))
    assert "out of range" in str(e.value)
    o, d = random_rays(64, 64, seed=1)
    rb = V.RayBatch.from_numpy(o, d)
    for bad in (0, 70000):
        with pytest.raises(V.VrtError):
            world.cast(rb, bad)
    # a misaligned hit record buffer is rejected
    h = V.HitBatch.empty(64)
    raw = torch.empty(64 * 16 + 8, dtype=torch.uint8, device="cuda")
    hits = L.Hits(h.t.data_ptr(), raw.data_ptr() + 4, 0)
    r = rb.to_c()
    import ctypes as C

    assert V.lib().vrt_cast(world.handle, C.byref(r), 128, C.byref(hits), None) == L.VRT_ERR_INVALID_ARGUMENT
    # zero rays and an empty sync are fine
    world.cast(V.RayBatch.from_numpy(np.zeros((3, 0), np.float32), np.zeros((3, 0), np.float32)), 128)
    world.sync(np.zeros(0, np.uint32), np.zeros(0, np.uint64), np.zeros(0, np.uint8))


def test_forced_hip_failure_is_reported(vrt, cuda):
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L

    world = V.World(6, 6)
    o, d = random_rays(64, 64, seed=1)
    rb = V.RayBatch.from_numpy(o, d)
    os.environ["VRT_FAULT_INJECT"] = "vrt_cast"
    try:
        with pytest.raises(V.VrtError) as e:
            world.cast(rb, 128)
        assert e.value.status == L.VRT_ERR_HIP and "injected" in str(e.value)
    finally:
        del os.environ["VRT_FAULT_INJECT"]
    world.cast(rb, 128)
    torch.cuda.synchronize()


@pytest.mark.parametrize("sparse", [True, False])
def test_c3_spheres_parity(vrt, oracle_mod, cuda, sparse):
    """2048^3 spheres (587 sparse / 20k literal, r ~ U(4, 64)): 1920x1080 primary rays plus
    2M rays with origins uniform in the box and directions uniform on the sphere."""
    import time

    import voxelrt_amd as V
    from voxelrt_amd import workloads as W

    scene = W.c3_world(sparse)
    world = V.World.from_scene(scene)
    ow = oracle_mod.OracleWorld.from_stream(scene.stream())
    cam = W.c3_camera()
    prim = cam.rays_device()
    o_i, d_i = V.incoherent_rays_host(2_000_000, 2048, 2048, seed=33)
    inc = V.RayBatch.from_numpy(o_i, d_i)
    for name, rb in (("primary", prim), ("incoherent", inc)):
        h = world.cast(rb, W.MAX_ITERS)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            world.cast(rb, W.MAX_ITERS, out=h)
        torch.cuda.synchronize()
        rate = rb.n * 5 / (time.perf_counter() - t0) / 1e6
        ref = ow.cast_simd(rb.o.cpu().numpy(), rb.d.cpu().numpy(), rb.world_origin, W.MAX_ITERS)
        assert_hits_equal(h.t.cpu().numpy(), h.rec.cpu().numpy(), ref["t"], ref["rec"], f"C3 {name}")
        # the compact schedule (first passes of 24, 32 and 48 steps) on the full sets, with the
        # grid's lookup, descent and hit totals
        ctr_g = V.new_counters()
        world.cast(rb, W.MAX_ITERS, counters=ctr_g)
        for cap in (24, 0, 48):
            ctr = V.new_counters()
            hc = world.cast(rb, W.MAX_ITERS, schedule=3, pass_cap=cap, counters=ctr)
            torch.cuda.synchronize()
            assert_hits_equal(hc.t.cpu().numpy(), hc.rec.cpu().numpy(), ref["t"], ref["rec"], f"C3 {name} cap={cap}")
            assert np.array_equal(hc.t.cpu().numpy().view(np.uint32), ref["t"].view(np.uint32))
            assert torch.equal(ctr, ctr_g), (cap, ctr.tolist(), ctr_g.tolist())
        hit = float(((ref["rec"][:, 3] & 0xFF) != 0).mean())
        print(f"C3 sparse={sparse} {name}: {rb.n} rays, {rate:.0f} Mrays/s, hit fraction {hit:.3f}", flush=True)


def test_maximum_extent_world(vrt, cuda):
    """The largest world (16384^3 voxels: 512^3 sectors, a 2 GiB sector table): one voxel near
    the far corner is found along each axis from across the world, at each world-origin
    specialization, at the closed-form distance to its entry face; a ray crossing the world
    away from it misses.  The sync of one sector updates that record only (scatter path)."""
    import voxelrt_amd as V

    n = 1 << 14
    world = V.World(14, 14)
    assert world.info().num_sectors == 1 << 27
    X, Y, Z = n - 4, n - 3, n - 2
    si = (X >> 5) | (Z >> 5) << 9 | (Y >> 5) << 18
    bi = ((X >> 3) & 3) | ((Z >> 3) & 3) << 2 | ((Y >> 3) & 3) << 4
    brick = np.zeros((1, 512), np.uint8)
    brick[0, (X & 7) | (Z & 7) << 3 | (Y & 7) << 6] = 9
    world.sync(np.array([si]), np.array([1 << bi], np.uint64), brick)
    e = 2.0 ** -20
    cases = [  # absolute origin, direction, closed-form t (None: a miss), normal
        ((X + 0.5, Y + 0.5, 0.5), (e, e, 1.0), Z - 0.5, (0, 0, -1)),
        ((X + 0.5, 0.5, Z + 0.5), (e, 1.0, e), Y - 0.5, (0, -1, 0)),
        ((0.5, Y + 0.5, Z + 0.5), (1.0, e, e), X - 0.5, (-1, 0, 0)),
        ((n - 0.25, Y + 0.5, Z + 0.5), (-1.0, e, e), n - 0.25 - (X + 1), (1, 0, 0)),
        ((0.5, 0.5, 0.5), (e, 1.0, e), None, None),
    ]
    absolute = np.array([c[0] for c in cases], np.float64).T
    d = np.array([c[1] for c in cases], np.float32).T
    # world origins of each specialization of the step: zero, a sector corner, any
    for W in [(0, 0, 0), (X & ~31, Y & ~31, 0), (X - 3, 7, 5)]:
        o = (absolute - np.array(W, np.float64)[:, None]).astype(np.float32)
        assert np.array_equal(o.astype(np.float64) + np.array(W, np.float64)[:, None], absolute)
        h = world.cast(V.RayBatch.from_numpy(o, d, W), 4096)
        torch.cuda.synchronize()
        t, rec = h.t.cpu().numpy(), h.rec.cpu().numpy()
        for k, (_, _, tk, nk) in enumerate(cases):
            info = int(rec[k, 3]) & 0xFFFFFFFF
            if tk is None:
                assert info & 0xFF == 0 and t[k] == -1.0, (W, k)
                continue
            normal = tuple(((info >> (8 + 2 * a)) & 3 ^ 2) - 2 for a in range(3))
            assert (tuple(rec[k, :3]), info & 0xFF, normal) == ((X, Y, Z), 9, nk), (W, k)
            assert t[k] == np.float32(tk), (W, k, t[k], tk)


def brick_alloc(vox: np.ndarray, si: int) -> tuple[int, list]:
    """Allocation mask of sector si of a dense [Y][Z][X] grid and its 64 bricks' bytes."""
    sy, sz, sx = vox.shape
    nxz = sx // 32
    sh = nxz.bit_length() - 1
    cx, cz, cy = si & (nxz - 1), (si >> sh) & (nxz - 1), si >> (2 * sh)
    sec = vox[cy * 32:(cy + 1) * 32, cz * 32:(cz + 1) * 32, cx * 32:(cx + 1) * 32]
    m, data = 0, []
    for b in range(64):
        bx, bz, by = b & 3, (b >> 2) & 3, b >> 4
        br = sec[by * 8:(by + 1) * 8, bz * 8:(bz + 1) * 8, bx * 8:(bx + 1) * 8].reshape(-1)
        data.append(br)
        if br.any():
            m |= 1 << b
    return m, data


def dirty_stream(vox: np.ndarray, dirty: dict):
    """(sector_idx, alloc_mask, dirty_mask, bricks) for {sector: dirty brick mask}: only the dirty
    bricks that stay allocated are sent, in increasing brick index (dirtyMask &= allocMask)."""
    idx, am, dm, bricks = [], [], [], []
    for si, dmask in dirty.items():
        m, data = brick_alloc(vox, si)
        idx.append(si)
        am.append(m)
        dm.append(dmask)
        bricks += [data[b] for b in range(64) if (dmask & m) >> b & 1]
    br = np.concatenate(bricks) if bricks else np.zeros(0, np.uint8)
    return np.array(idx, np.uint32), np.array(am, np.uint64), np.array(dm, np.uint64), br


def brick_of(x: int, y: int, z: int, nxz_log2: int = 2) -> tuple[int, int]:
    si = (x >> 5) | (z >> 5) << nxz_log2 | (y >> 5) << (2 * nxz_log2)
    bi = ((x >> 3) & 3) | ((z >> 3) & 3) << 2 | ((y >> 3) & 3) << 4
    return si, bi


def test_brick_sync_uploads_only_dirty_bricks(vrt, oracle_mod, cuda):
    """DirtyLocs at brick granularity (GpuRenderer.cpp:52-79, 114-129): single-brick edits of full
    sectors upload 512 bytes each; allocation changes move the kept bricks on the device; after
    every edit the casts equal the oracle's on the edited grid and a freshly uploaded world's."""
    import voxelrt_amd as V
    from test_oracle import dense_world

    rng = np.random.default_rng(17)
    vox = blob_world(128, seed=41, density=0.02)
    vox[:40] = np.where(vox[:40] == 0, 7, vox[:40])  # solid floor: sectors of the bottom layer are full
    vox[16:24, 8:16, 72:80] = 0                       # one empty brick (37 of sector 2) inside a full sector
    nsec = 64
    world = V.World(7, 7)
    world.sync(*sector_stream_of(vox, list(range(nsec))))
    o, d = random_rays(20000, 128, seed=19)

    def check(tag):
        ref = dense_world(oracle_mod, vox).cast(o, d, (0, 0, 0), 128)
        t, rec = cast_np(world, o, d)
        assert_hits_equal(t, rec, ref["t"], ref["rec"], tag)
        fresh = V.World.from_scene(V.Scene.from_dense(vox))
        assert np.array_equal(rec, cast_np(fresh, o, d)[1]), tag
        masks = sector_stream_of(vox, list(range(nsec)))[1]
        check_sector_levels(world, dict(zip(range(nsec), masks)))
        assert world.info().num_bricks == sum(bin(int(m)).count("1") for m in masks), tag

    check("initial")
    full_si, _ = brick_of(4, 4, 4)
    assert brick_alloc(vox, full_si)[0] == (1 << 64) - 1
    assert brick_alloc(vox, 2)[0] == ((1 << 64) - 1) & ~(1 << 37)

    def sync(dirty, tag, bricks_up, moved):
        nonlocal vox
        i0 = world.info()
        world.sync_bricks(*dirty_stream(vox, dirty))
        i1 = world.info()
        assert i1.upload_brick_bytes - i0.upload_brick_bytes == 512 * bricks_up, tag
        assert (i1.moved_bricks - i0.moved_bricks > 0) == moved, tag
        # the brick bytes and their slots, the records and the changed sector-of-sectors words
        assert i1.upload_bytes - i0.upload_bytes <= 516 * bricks_up + 8 * (i1.moved_bricks - i0.moved_bricks) \
            + 48 * len(dirty) + 64, tag
        check(tag)

    # 1. voxels repainted inside one brick of a full sector: that brick alone
    vox[9, 10, 11] = 200
    vox[12, 13, 14] = 0
    sync({brick_of(11, 9, 10)[0]: 1 << brick_of(11, 9, 10)[1]}, "repaint", 1, False)
    # 2. a voxel in the empty brick: the allocation grows in the middle, kept bricks shift
    vox[22, 10, 75] = 5
    si, bi = brick_of(75, 22, 10)
    sync({si: 1 << bi}, "new brick", 1, True)
    # 3. a brick cleared: it leaves the allocation, nothing uploaded, the bricks after it shift
    vox[0:8, 0:8, 0:8] = 0
    si, bi = brick_of(0, 0, 0)
    sync({si: 1 << bi}, "cleared brick", 0, True)
    # 4. several sectors in one sync, random single voxels, dirty masks listing some clean bricks
    dirty = {}
    for _ in range(12):
        x, y, z = (int(c) for c in rng.integers(0, 128, 3))
        vox[y, z, x] = int(rng.integers(0, 256))
        s, b = brick_of(x, y, z)
        dirty[s] = dirty.get(s, 0) | 1 << b
    s0 = next(iter(dirty))
    dirty[s0] |= 1 << 63  # a clean brick listed as dirty is sent again (if allocated)
    n_up = sum(bin(d & brick_alloc(vox, s)[0]).count("1") for s, d in dirty.items())
    i0 = world.info()
    world.sync_bricks(*dirty_stream(vox, dirty))
    assert world.info().upload_brick_bytes - i0.upload_brick_bytes == 512 * n_up
    check("scattered edits")
    # 5. a whole sector emptied
    vox[32:64, 32:64, 32:64] = 0
    si = 1 | 1 << 2 | 1 << 4
    sync({si: (1 << 64) - 1}, "emptied sector", 0, False)
    # a brick newly allocated but not dirty is rejected and leaves the world unchanged
    vox2 = vox.copy()
    vox2[40, 40, 40] = 3
    si, bi = brick_of(40, 40, 40)
    idx, am, dm, br = dirty_stream(vox2, {si: 0})
    with pytest.raises(V.VrtError, match="not dirty"):
        world.sync_bricks(idx, am, dm, br)
    check("rejected")


def test_scene_edits_sync_through_the_dirty_set(vrt, oracle_mod, cuda):
    """The reference's editing loop: VoxelMap::Set marks DirtyLocs, SyncBuffers(VoxelMap&)
    consumes it (vrt_scene_set_voxel, vrt_world_sync_scene).  Each round of edits uploads
    exactly the dirty bricks still allocated, clears the dirty set, and leaves the world's casts
    equal to the oracle's over the scene's sector stream (allocated empty bricks included), every
    hit on a solid voxel of the edited dense grid."""
    import voxelrt_amd as V

    rng = np.random.default_rng(23)
    vox = blob_world(128, seed=43, density=0.01)
    scene = V.Scene.from_dense(vox)
    world = V.World.from_scene(scene)
    o, d = random_rays(20000, 128, seed=24)
    for rnd in range(4):
        dirty = {}
        for _ in range(300 if rnd < 3 else 0):
            x, y, z = (int(c) for c in rng.integers(0, 128, 3))
            v = int(rng.integers(0, 256)) if rng.random() < 0.8 else 0
            scene.set_voxel(x, y, z, v)
            vox[y, z, x] = v
            s, b = brick_of(x, y, z)
            dirty[s] = dirty.get(s, 0) | 1 << b
        if rnd == 3:
            # carve a whole sector empty, then collect the empty bricks
            for y in range(64, 96):
                for z in range(0, 32):
                    for x in range(32, 64):
                        if vox[y, z, x]:
                            scene.set_voxel(x, y, z, 0)
                            vox[y, z, x] = 0
                            s, b = brick_of(x, y, z)
                            dirty[s] = dirty.get(s, 0) | 1 << b
            scene.delete_empty_bricks()
        assert scene.dirty_sectors >= len(dirty)
        # every dirty brick that is still allocated goes up (Set allocates even for 0)
        st = scene.stream()
        present = dict(zip(st["sector_idx"].tolist(), st["alloc_mask"].tolist()))
        alloc = {s: int(present.get(s, 0)) for s in dirty}
        n_up = sum(bin(dm & alloc[s]).count("1") for s, dm in dirty.items())
        i0 = world.info()
        world.sync_scene(scene)
        assert scene.dirty_sectors == 0
        i1 = world.info()
        assert i1.upload_brick_bytes - i0.upload_brick_bytes == 512 * n_up, f"round {rnd}"
        # the oracle over the scene's own sector stream: Set allocates bricks also for 0, and an
        # allocated empty brick changes the steps' LOD (not the hits), as in the reference
        ref = oracle_mod.OracleWorld.from_stream(scene.stream()).cast(o, d, (0, 0, 0), 128)
        t, rec = cast_np(world, o, d)
        assert_hits_equal(t, rec, ref["t"], ref["rec"], f"round {rnd}")
        # every hit is a solid voxel of the edited grid, with its material
        hit = (rec[:, 3] & 0xFF) != 0
        hx, hy, hz = rec[hit, 0], rec[hit, 1], rec[hit, 2]
        assert hit.any() and np.array_equal(vox[hy, hz, hx], (rec[hit, 3] & 0xFF).astype(np.uint8))
        assert world.info().num_bricks == scene.stream()["nbricks"]


def test_mark_all_dirty_fills_an_empty_world(vrt, oracle_mod, cuda):
    """MarkAllDirty then SyncBuffers into an empty world (the reference's re-upload after a buffer
    resize, GpuRenderer.cpp:85-97) gives the world vrt_world_from_scene builds; a world of other
    extents is refused and the dirty set kept; the scene's palette, once set, goes up with each
    sync (GpuRenderer.cpp:100-102), and a scene without one leaves the world's."""
    import voxelrt_amd as V

    vox = blob_world(128, seed=44, density=0.01)
    scene = V.Scene.from_dense(vox)
    scene.mark_all_dirty()
    assert scene.dirty_sectors == len(scene.stream()["sector_idx"])
    with pytest.raises(V.VrtError, match="extents"):
        V.World(6, 6).sync_scene(scene)
    assert scene.dirty_sectors > 0
    world = V.World(7, 7)
    world.sync_scene(scene)
    assert scene.dirty_sectors == 0
    st = scene.stream()
    assert world.info().num_bricks == st["nbricks"]
    assert world.info().upload_brick_bytes == 512 * st["nbricks"]
    o, d = random_rays(20000, 128, seed=45)
    t, rec = cast_np(world, o, d)
    t2, rec2 = cast_np(V.World.from_scene(scene), o, d)
    assert np.array_equal(rec, rec2) and np.array_equal(t.view(np.uint32), t2.view(np.uint32))
    ref = oracle_mod.OracleWorld.from_stream(st).cast(o, d, (0, 0, 0), 128)
    assert_hits_equal(t, rec, ref["t"], ref["rec"], "mark all dirty")

    # palette: a scene without one leaves the world's; a scene's palette goes up with each sync
    def material_data():
        h = world.cast(V.RayBatch.from_numpy(o, d), 128, with_material=True)
        torch.cuda.synchronize()
        return h.rec[:, 3].cpu().numpy() & 0xFF, h.material_data.cpu().numpy().astype(np.uint32)

    mine = (np.arange(256, dtype=np.uint64) * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)
    world.set_palette(mine)
    world.sync_scene(scene)
    mat, md = material_data()
    assert np.array_equal(md[mat != 0], mine[mat[mat != 0]].astype(np.uint32))
    theirs = (np.arange(256, dtype=np.uint64) * np.uint64(0x85EBCA6B) + np.uint64(7)) & np.uint64(0xFFFFFFFF)
    scene.set_palette(theirs)
    world.sync_scene(scene)
    mat, md = material_data()
    assert np.array_equal(md[mat != 0], theirs[mat[mat != 0]].astype(np.uint32))


def _device_bytes(ptr: int, nbytes: int) -> np.ndarray:
    torch.cuda.synchronize()
    return torch.as_tensor(_DeviceWords(ptr, (nbytes // 4,)), device="cuda").cpu().numpy().view(np.uint8)


def _cell_masks(brick: np.ndarray) -> list[int]:
    """The 8 occupancy masks of a brick [y][z][x] (8^3 bytes), by brute force: cell cx + 2cz + 4cy,
    bit vx + 4vz + 16vy (CpuRenderer.cpp:63-83)."""
    out = []
    for cy in range(2):
        for cz in range(2):
            for cx in range(2):
                m = 0
                for vy in range(4):
                    for vz in range(4):
                        for vx in range(4):
                            if brick[cy * 4 + vy, cz * 4 + vz, cx * 4 + vx]:
                                m |= 1 << (vx + 4 * vz + 16 * vy)
                out.append(m)  # appended in cell order cx + 2cz + 4cy
    return out


def test_device_layout_matches_brute_force(vrt, cuda):
    """The device world read back after incremental syncs that move bricks: every allocated brick
    sits in the slot an explicit per-brick slot table gives (bricks of a sector in increasing
    index from its base slot, BrickSlotAllocator.h:37-41), holds its voxel bytes, and its 8
    occupancy masks equal a brute-force build from those bytes; masks split at bit 32 included."""
    import voxelrt_amd as V

    rng = np.random.default_rng(29)
    vox = blob_world(128, seed=47, density=0.01)
    vox[64:96, 0:32, 0:32] = 0
    vox[80:96, 0:32, 0:32] = np.where(rng.random((16, 32, 32)) < 0.05, 9, 0)  # bricks 32..63 only
    vox[0:32, 96:128, 96:128] = 0
    vox[0:16, 96:128, 96:128] = np.where(rng.random((16, 32, 32)) < 0.05, 11, 0)  # bricks 0..31 only
    world = V.World(7, 7)
    world.sync(*sector_stream_of(vox, list(range(64))))
    split_sectors = {(0, 2, 0), (3, 0, 3)}  # (x, y, z) of the two sectors filled above: left alone
    for _ in range(3):  # edits that grow, shrink and move sectors' slot runs
        done = 0
        while done < 6:
            x, y, z = (int(c) for c in rng.integers(0, 120, 3))
            touched = {(a >> 5, b >> 5, c >> 5) for a in (x, x + 7) for b in (y, y + 7) for c in (z, z + 7)}
            if touched & split_sectors:
                continue
            before = vox.copy()
            vox[y:y + 8, z:z + 8, x:x + 8] = 0 if rng.random() < 0.5 else int(rng.integers(1, 256))
            done += 1
            # only the bricks that changed are dirty, so the bricks kept around them move when
            # their sector's slot run changes
            dirty = {}
            for bz in range(z >> 3, ((z + 7) >> 3) + 1):
                for by in range(y >> 3, ((y + 7) >> 3) + 1):
                    for bx in range(x >> 3, ((x + 7) >> 3) + 1):
                        sl = np.s_[by * 8:by * 8 + 8, bz * 8:bz * 8 + 8, bx * 8:bx * 8 + 8]
                        if not np.array_equal(before[sl], vox[sl]):
                            si, bi = brick_of(bx * 8, by * 8, bz * 8)
                            dirty[si] = dirty.get(si, 0) | 1 << bi
            if dirty:
                world.sync_bricks(*dirty_stream(vox, dirty))
    info = world.info()
    buf = world.buffers()
    rec = sector_records(world)
    occ = _device_bytes(buf.occupancy, info.slot_capacity * 64).view(np.uint64)
    dev_vox = _device_bytes(buf.voxels, info.slot_capacity * 512)
    used = set()
    split = {"lo": 0, "hi": 0}
    for si in range(64):
        mask = int(rec[si, 0]) | int(rec[si, 1]) << 32
        base = int(rec[si, 2])
        sx, sz, sy = si & 3, (si >> 2) & 3, si >> 4
        sec = vox[sy * 32:(sy + 1) * 32, sz * 32:(sz + 1) * 32, sx * 32:(sx + 1) * 32]
        slot = base
        for b in range(64):
            bx, bz, by = b & 3, (b >> 2) & 3, b >> 4
            brick = sec[by * 8:(by + 1) * 8, bz * 8:(bz + 1) * 8, bx * 8:(bx + 1) * 8]
            if not (mask >> b) & 1:
                assert not brick.any(), (si, b)
                continue
            assert slot not in used, (si, b, slot)
            used.add(slot)
            assert np.array_equal(dev_vox[slot * 512:(slot + 1) * 512], brick.reshape(-1)), (si, b)
            assert [int(m) for m in occ[slot * 8:(slot + 1) * 8]] == _cell_masks(brick), (si, b)
            slot += 1
        if mask and mask >> 32 == 0:
            split["lo"] += 1
        if mask and mask & 0xFFFFFFFF == 0:
            split["hi"] += 1
    assert split["lo"] > 0 and split["hi"] > 0 and world.info().moved_bricks > 0
    assert len(used) == info.num_bricks


def test_sync_staging_failure_leaves_world_unchanged(vrt, cuda):
    """The staging buffers are allocated by a world's first sync, before any device work: a failure
    there returns an error and the world stays as it was (empty), usable by the next sync."""
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L

    vox = blob_world(64, seed=8, density=0.01)
    o, d = random_rays(4000, 64, seed=2)
    world = V.World(6, 6)
    os.environ["VRT_FAULT_INJECT"] = "vrt_world_sync:staging"
    try:
        with pytest.raises(V.VrtError) as e:
            world.sync(*sector_stream_of(vox, list(range(8))))
        assert e.value.status == L.VRT_ERR_HIP
    finally:
        del os.environ["VRT_FAULT_INJECT"]
    assert world.info().num_bricks == 0
    t, rec = cast_np(world, o, d)
    assert np.all((rec[:, 3] & 0xFF) == 0)
    world.sync(*sector_stream_of(vox, list(range(8))))
    fresh = V.World.from_scene(V.Scene.from_dense(vox))
    assert np.array_equal(cast_np(world, o, d)[1], cast_np(fresh, o, d)[1])


def test_sync_device_failure_marks_world_failed(vrt, cuda):
    """A HIP failure once a sync's device work has begun leaves the device tables part-written:
    the call reports it, and every later call on the world returns VRT_ERR_STATE."""
    import voxelrt_amd as V
    from voxelrt_amd import _lib as L

    vox = blob_world(64, seed=8, density=0.01)
    world = V.World(6, 6)
    world.sync(*sector_stream_of(vox, list(range(8))))
    o, d = random_rays(64, 64, seed=1)
    rb = V.RayBatch.from_numpy(o, d)
    vox[1, 2, 3] = 9
    os.environ["VRT_FAULT_INJECT"] = "vrt_world_sync:device"
    try:
        with pytest.raises(V.VrtError) as e:
            world.sync(*sector_stream_of(vox, [0]))
        assert e.value.status == L.VRT_ERR_HIP and "unusable" in str(e.value)
    finally:
        del os.environ["VRT_FAULT_INJECT"]
    torch.cuda.synchronize()
    for call in (lambda: world.cast(rb, 128), lambda: world.sync(*sector_stream_of(vox, [0])),
                 lambda: world.set_palette(np.zeros(256, np.uint64))):
        with pytest.raises(V.VrtError) as e:
            call()
        assert e.value.status == L.VRT_ERR_STATE and "destroy" in str(e.value)
