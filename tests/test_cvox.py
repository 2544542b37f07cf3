"""The reference's world file ("cvox 0004", VoxelMap.cpp:207-274) through vrt_scene_load_cvox /
vrt_scene_save_cvox.  No file of this format ships with the reference, so files are written here
by an independent Python writer that follows the layout (zstd blobs of BinaryIO.cpp:8-60, made
with pyarrow's zstd codec), and by the library's own writer for round trips."""
from __future__ import annotations

import struct

import numpy as np
import pytest

pa = pytest.importorskip("pyarrow")

MAGIC = 0x0000000478766F63


def blob(raw: bytes) -> bytes:
    comp = pa.compress(raw, codec="zstd", asbytes=True)
    return struct.pack("<I", len(comp)) + comp


def world_index(x: int, y: int, z: int) -> int:
    """WorldSectorIndexer<12,8,signed>::GetIndex (VoxelMap.h:95-97)."""
    return (x & 0xFFF) | (z & 0xFFF) << 12 | (y & 0xFF) << 24


def material_records(rng, n=256):
    rec = np.zeros((n, 16), np.uint8)
    rec[:, :4] = rng.integers(0, 256, (n, 4))
    em = rng.uniform(0, 8, n).astype(np.float32)
    em[::7] = 0.0
    em[3] = np.float32(1.0009765625)   # exactly between two halves: ties to even
    em[5] = np.float32(70000.0)        # above the half range: infinity
    em[6] = np.float32(3e-6)           # half subnormal
    rec[:, 4:8] = em.view(np.uint8).reshape(n, 4)
    return rec, em


def write_file(path, sectors, materials, pack_limit=None):
    """sectors: list of (world sector (x, y, z), mask, bricks uint8[popcount * 512])."""
    out = [struct.pack("<QI", MAGIC, len(sectors)), blob(materials.tobytes())]
    packs, cur = [], b""
    for (x, y, z), mask, bricks in sectors:
        cur += struct.pack("<IQ", world_index(x, y, z), mask) + bricks.tobytes()
        if pack_limit is not None and len(cur) >= pack_limit:
            packs.append(cur)
            cur = b""
    packs.append(cur)
    for pk in packs:
        out.append(struct.pack("<I", len(pk)) + blob(pk))
    with open(path, "wb") as f:
        f.write(b"".join(out))


def random_sectors(rng, coords):
    secs = []
    for c in coords:
        mask = int(rng.integers(1, 1 << 63, dtype=np.uint64)) | 1
        nb = bin(mask).count("1")
        bricks = rng.integers(0, 256, nb * 512).astype(np.uint8)
        bricks[rng.random(nb * 512) < 0.6] = 0
        secs.append((c, mask, bricks))
    return secs


def test_load_independent_file(vrt, tmp_path):
    rng = np.random.default_rng(1)
    coords = [(0, 0, 0), (3, 1, 2), (-1, 0, 0), (1, 2, 3), (5, 0, 0), (2, -3, 1)]
    secs = random_sectors(rng, coords)
    mats, em = material_records(rng)
    path = tmp_path / "w.cvox"
    write_file(path, secs, mats, pack_limit=20000)  # several packs
    sc, info = vrt.Scene.load_cvox(str(path), 7, 7, sector_offset=(-1, -1, 0))
    # scene extents: 4 sectors per axis; world sector s lands at s - offset = s + (1, 1, 0)
    assert info["sectors_in_file"] == 6
    inside = [(c, m, b) for c, m, b in secs if 0 <= c[0] + 1 < 4 and 0 <= c[1] + 1 < 4 and 0 <= c[2] < 4]
    assert info["sectors_loaded"] == len(inside) and info["sectors_outside"] == 6 - len(inside)
    for (x, y, z), mask, bricks in inside:
        sx, sy, sz = x + 1, y + 1, z
        k = 0
        for b in range(64):
            if not (mask >> b) & 1:
                continue
            brick = bricks[k * 512:(k + 1) * 512]
            k += 1
            for v in (0, 1, 77, 511):
                vx, vz, vy = v & 7, (v >> 3) & 7, v >> 6
                bx, bz, by = b & 3, (b >> 2) & 3, b >> 4
                gx, gy, gz = sx * 32 + bx * 8 + vx, sy * 32 + by * 8 + vy, sz * 32 + bz * 8 + vz
                assert sc.voxel(gx, gy, gz) == brick[v]
    pal = sc.palette()
    with np.errstate(over="ignore"):
        half = em.astype(np.float16).view(np.uint16).astype(np.uint64)
    want = ((mats[:, 0].astype(np.uint64) >> 3) << 11) | ((mats[:, 1].astype(np.uint64) >> 2) << 5) | \
        (mats[:, 2].astype(np.uint64) >> 3) | (half << 16) | (mats[:, 3].astype(np.uint64) << 32)
    assert np.array_equal(pal, want)


def test_round_trip_and_errors(vrt, tmp_path):
    from helpers import blob_world

    vox = blob_world(128, seed=3, density=0.01, blobs=4)
    sc = vrt.Scene.from_dense(vox)
    rng = np.random.default_rng(2)
    mats, _ = material_records(rng)
    path = tmp_path / "rt.cvox"
    sc.save_cvox(str(path), sector_offset=(-2, 0, 5), materials=mats)
    back, info = vrt.Scene.load_cvox(str(path), 7, 7, sector_offset=(-2, 0, 5))
    a, b = sc.stream(), back.stream()
    for k in ("sector_idx", "alloc_mask", "bricks"):
        assert np.array_equal(a[k], b[k]), k
    assert info["bricks_loaded"] == a["nbricks"] and b["solid_voxels"] == int((vox != 0).sum())
    # the library's file is readable by the independent reader's layout: magic, count, blobs
    raw = path.read_bytes()
    magic, n = struct.unpack_from("<QI", raw)
    assert magic == MAGIC and n == a["nsec"]
    clen = struct.unpack_from("<I", raw, 12)[0]
    assert pa.decompress(raw[16:16 + clen], decompressed_size=4096, codec="zstd", asbytes=True) == mats.tobytes()
    # errors: wrong magic, truncation, a blob that inflates to the wrong size
    bad = tmp_path / "bad.cvox"
    bad.write_bytes(struct.pack("<QI", MAGIC + 1, 0))
    with pytest.raises(vrt.VrtError, match="not a cvox"):
        vrt.Scene.load_cvox(str(bad), 7, 7)
    bad.write_bytes(raw[: len(raw) // 2])
    with pytest.raises(vrt.VrtError):
        vrt.Scene.load_cvox(str(bad), 7, 7, sector_offset=(-2, 0, 5))
    bad.write_bytes(struct.pack("<QI", MAGIC, 0) + blob(b"\0" * 100))
    with pytest.raises(vrt.VrtError, match="fewer bytes"):
        vrt.Scene.load_cvox(str(bad), 7, 7)
    with pytest.raises(vrt.VrtError, match="cannot open"):
        vrt.Scene.load_cvox(str(tmp_path / "missing.cvox"), 7, 7)


def test_committed_fixture_loads_to_its_grid(vrt):
    """tests/golden/terrain128.cvox (written by tests/golden/make_cvox.py) loads to exactly the
    dense grid it was written from, with the palette of its material records."""
    from pathlib import Path

    g = Path(__file__).resolve().parent / "golden"
    f = np.load(g / "terrain128.npz")
    sc, info = vrt.Scene.load_cvox(str(g / "terrain128.cvox"), 7, 7,
                                   sector_offset=tuple(int(x) for x in f["sector_offset"]))
    ref = vrt.Scene.from_dense(f["voxels"])
    a, b = sc.stream(), ref.stream()
    for k in ("sector_idx", "alloc_mask", "bricks"):
        assert np.array_equal(a[k], b[k]), k
    mats = f["materials"]
    em = mats[:, 4:8].copy().view(np.float32).ravel()
    with np.errstate(over="ignore"):
        half = em.astype(np.float16).view(np.uint16).astype(np.uint64)
    want = ((mats[:, 0].astype(np.uint64) >> 3) << 11) | ((mats[:, 1].astype(np.uint64) >> 2) << 5) | \
        (mats[:, 2].astype(np.uint64) >> 3) | (half << 16) | (mats[:, 3].astype(np.uint64) << 32)
    assert np.array_equal(sc.palette(), want)
