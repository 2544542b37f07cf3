"""Writes tests/golden/terrain128.cvox and tests/golden/terrain128.npz.

The reference ships no world file of its format ("cvox 0004", VoxelMap.cpp:207-274), so this
script writes one with the independent writer of tests/test_cvox.py (the layout of
VoxelMap::Serialize and BinaryIO.cpp:8-60, zstd blobs made with pyarrow's codec), from a
seeded 128^3 terrain with caves and height-band materials.  terrain128.npz keeps the dense grid
the file was written from ([y][z][x] u8), the world sector offset the file uses and the
material records, so a test can load the file through the library and compare with a world
built from the grid itself.
Run from the repository root: python tests/golden/make_cvox.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from test_cvox import material_records, write_file  # noqa: E402

# world sectors of the file = scene sectors + this offset (negative coordinates on purpose:
# WorldSectorIndexer<12,8,signed> wraps them, VoxelMap.h:95-100)
SECTOR_OFFSET = (-2, 1, -1)


def terrain(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    # smooth height field: a sum of a few random cosine waves
    x = np.arange(n)[None, :].astype(np.float64)
    z = np.arange(n)[:, None].astype(np.float64)
    h = np.full((n, n), 40.0)
    for k in range(6):
        fx, fz = rng.uniform(0.01, 0.08, 2)
        ph = rng.uniform(0, 2 * np.pi)
        h += rng.uniform(4, 14) * np.cos(fx * x * 2 * np.pi / 8 + fz * z * 2 * np.pi / 8 + ph)
    y = np.arange(n)[:, None, None]
    solid = y < h[None, :, :]
    # caves: a few ellipsoids carved out
    yy, zz, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for _ in range(5):
        c = rng.uniform(10, n - 10, 3)
        r = rng.uniform(4, 10, 3)
        solid &= ((xx - c[0]) / r[0]) ** 2 + ((yy - c[1] / 3) / r[1]) ** 2 + ((zz - c[2]) / r[2]) ** 2 > 1.0
    mat = np.clip(1 + yy // 8, 1, 255).astype(np.uint8)
    return np.where(solid, mat, 0).astype(np.uint8)


def sectors_of(vox: np.ndarray):
    n = vox.shape[0]
    out = []
    for sy in range(n // 32):
        for sz in range(n // 32):
            for sx in range(n // 32):
                sec = vox[sy * 32:(sy + 1) * 32, sz * 32:(sz + 1) * 32, sx * 32:(sx + 1) * 32]
                mask, bricks = 0, []
                for b in range(64):
                    bx, bz, by = b & 3, (b >> 2) & 3, b >> 4
                    br = sec[by * 8:(by + 1) * 8, bz * 8:(bz + 1) * 8, bx * 8:(bx + 1) * 8]
                    if br.any():
                        mask |= 1 << b
                        bricks.append(br.reshape(-1))
                if mask:
                    w = (sx + SECTOR_OFFSET[0], sy + SECTOR_OFFSET[1], sz + SECTOR_OFFSET[2])
                    out.append((w, mask, np.concatenate(bricks)))
    return out


def main():
    vox = terrain(128, seed=128)
    mats, _ = material_records(np.random.default_rng(129))
    secs = sectors_of(vox)
    write_file(HERE / "terrain128.cvox", secs, mats, pack_limit=1 << 20)
    np.savez_compressed(HERE / "terrain128.npz", voxels=vox, sector_offset=np.array(SECTOR_OFFSET, np.int32),
                        materials=mats)
    print(f"wrote terrain128.cvox: {len(secs)} sectors, {sum(bin(m).count('1') for _, m, _ in secs)} bricks, "
          f"{int((vox != 0).sum())} solid voxels")


if __name__ == "__main__":
    main()
