"""Test helpers: small seeded worlds and parity comparisons."""
from __future__ import annotations

import numpy as np


def blob_world(size: int, seed: int, density: float = 0.02, blobs: int = 6, ysize: int | None = None) -> np.ndarray:
    """Dense uint8 world [Y][Z][X]: random single voxels plus solid spheres, materials 1..255."""
    ysize = ysize or size
    rng = np.random.default_rng(seed)
    w = np.zeros((ysize, size, size), dtype=np.uint8)
    m = rng.random(w.shape) < density
    w[m] = rng.integers(1, 256, size=int(m.sum()), dtype=np.uint8)
    yy, zz, xx = np.meshgrid(np.arange(ysize), np.arange(size), np.arange(size), indexing="ij")
    for _ in range(blobs):
        c = rng.uniform(0, [ysize, size, size])
        r = rng.uniform(2, size / 6)
        inside = (yy + 0.5 - c[0]) ** 2 + (zz + 0.5 - c[1]) ** 2 + (xx + 0.5 - c[2]) ** 2 <= r * r
        w[inside] = rng.integers(1, 256, dtype=np.uint8)
    return w


def random_rays(n: int, size: int, seed: int, ysize: int | None = None, inside: bool = True):
    """Origins uniform in the world box (or a margin around it), directions uniform on the sphere."""
    ysize = ysize or size
    rng = np.random.default_rng(seed)
    lo = np.array([0, 0, 0]) if inside else np.array([-8, -8, -8])
    hi = np.array([size, ysize, size]) if inside else np.array([size + 8, ysize + 8, size + 8])
    o = rng.uniform(lo, hi, size=(n, 3)).T.astype(np.float32)
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    d = v.astype(np.float32)
    d[d == 0] = np.float32(2.0**-24)
    return np.ascontiguousarray(o), np.ascontiguousarray(d)


def assert_hits_equal(t_a, rec_a, t_b, rec_b, what: str = "") -> None:
    """Voxel, material, normal and iteration count bit-exact; t within 1e-5 relative (expected bit-exact)."""
    rec_a = np.asarray(rec_a)
    rec_b = np.asarray(rec_b)
    bad = np.nonzero(np.any(rec_a != rec_b, axis=1))[0]
    if bad.size:
        i = bad[0]
        raise AssertionError(f"{what}: {bad.size} hit records differ; first at ray {i}: {rec_a[i]} vs {rec_b[i]}")
    t_a = np.asarray(t_a, dtype=np.float32)
    t_b = np.asarray(t_b, dtype=np.float32)
    hit = (rec_a[:, 3] & 0xFF) != 0
    err = np.abs(t_a[hit].astype(np.float64) - t_b[hit].astype(np.float64))
    tol = 1e-5 * np.maximum(np.abs(t_b[hit].astype(np.float64)), 1e-30)
    bad_t = np.nonzero(err > tol)[0]
    if bad_t.size:
        i = np.nonzero(hit)[0][bad_t[0]]
        raise AssertionError(f"{what}: {bad_t.size} distances beyond 1e-5 relative; ray {i}: {t_a[i]} vs {t_b[i]}")
    assert np.all(t_a[~hit] == -1.0) and np.all(t_b[~hit] == -1.0), f"{what}: misses must report t = -1"
