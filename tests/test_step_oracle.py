"""Exhaustive pin of the oracle's GetStepPos (CpuRenderer.cpp:135-171) and of the
CPU method's iteration sequence, plus PCG known answers (RandomGen.glsl:6-20).  CPU only.

The closed form a step must meet: at a voxel whose brick is not allocated, the step
skips the largest empty aligned cube of bricks (edge 1, 2 or 4 bricks: 8, 16 or 32
voxels) that holds the voxel; in an allocated brick, a solid voxel is a hit and an
empty one skips the largest empty aligned cube of voxels (edge 1, 2 or 4) inside
its 4x4x4 cell.  The step lands on the cube's far corner in the ray's direction:
per axis the low corner when d < 0, else the high corner (so d = -0.0 goes high).
The expected values here are computed from the dense voxel grid and the allocation
bits by enumerating cubes in 3D, independently of the oracle's mask arithmetic
(32-bit halves, the 0xA group shift and the 0x00330033 group mask).

The planted-mistake tests compile a copy of the oracle with one mistake each and
check that the exhaustive comparison fails for it.
"""
from __future__ import annotations

import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"


# ---------------------------------------------------------------- scenario worlds


def brick_slice(b: int):
    bx, bz, by = b & 3, (b >> 2) & 3, b >> 4
    return np.s_[by * 8:(by + 1) * 8, bz * 8:(bz + 1) * 8, bx * 8:(bx + 1) * 8]


def scenario(kind: str, seed: int):
    """(alloc mask, dense 32^3 [y][z][x] voxels) of one sector.  Allocated bricks may be empty;
    unallocated bricks are empty."""
    rng = np.random.default_rng(seed)
    vox = np.zeros((32, 32, 32), np.uint8)
    if kind == "random":
        mask = int(rng.integers(0, 2**63)) | int(rng.integers(0, 2)) << 63
    elif kind == "low_half":        # bricks only in layers y 0..1 (mask bits 0..31)
        mask = int(rng.integers(1, 2**32))
    elif kind == "high_half":       # bricks only in layers y 2..3 (mask bits 32..63)
        mask = int(rng.integers(1, 2**32)) << 32
    elif kind == "empty":
        mask = 0
    elif kind == "full":
        mask = (1 << 64) - 1
    elif kind.startswith("single"):
        mask = 1 << int(kind[6:])
    else:
        raise ValueError(kind)
    for b in range(64):
        if not (mask >> b) & 1:
            continue
        style = int(rng.integers(0, 7))
        br = np.zeros((8, 8, 8), np.uint8)
        if style == 0:
            pass                                              # allocated, all empty
        elif style == 1:
            br[:] = 5                                         # full
        elif style == 2:
            br[:, :, :] = 0
            br[2:4, :, :] = 6                                 # cell layers y 2..3: each cell mask's high half only
            br[6:8, :, :] = 6
        elif style == 3:
            br[0:2, :, :] = 7                                 # cell layers y 0..1: low halves only
            br[4:6, :, :] = 7
        elif style == 4:                                      # one voxel in some cells
            for c in range(8):
                if rng.random() < 0.6:
                    cx, cz, cy = (c & 1) * 4, ((c >> 1) & 1) * 4, (c >> 2) * 4
                    p = rng.integers(0, 4, 3)
                    br[cy + p[1], cz + p[2], cx + p[0]] = 9
        else:
            dens = 0.05 if style == 5 else 0.5
            br[:] = np.where(rng.random((8, 8, 8)) < dens, rng.integers(1, 256, (8, 8, 8)), 0)
        vox[brick_slice(b)] = br
    return mask, vox


SCENARIOS = [("random", 1), ("random", 2), ("random", 3), ("low_half", 4), ("high_half", 5), ("empty", 6),
             ("full", 7), ("full", 8), ("single0", 9), ("single31", 10), ("single32", 11), ("single63", 12),
             ("single21", 13)]


# ---------------------------------------------------------------- closed form


def expected_steps(mask: int, vox: np.ndarray):
    """Per voxel (x, y, z) of the sector: solid, descended, and the edge E (voxels) of the
    empty aligned cube the step skips (0 on a hit)."""
    alloc = np.zeros((4, 4, 4), bool)  # [by][bz][bx]
    for b in range(64):
        alloc[b >> 4, (b >> 2) & 3, b & 3] = (mask >> b) & 1
    solid = np.zeros((32, 32, 32), bool)
    desc = np.zeros((32, 32, 32), bool)
    edge = np.zeros((32, 32, 32), np.int32)
    occ = vox != 0
    for by in range(4):
        for bz in range(4):
            for bx in range(4):
                sl = np.s_[by * 8:(by + 1) * 8, bz * 8:(bz + 1) * 8, bx * 8:(bx + 1) * 8]
                if not alloc[by, bz, bx]:
                    g = alloc[by & ~1:(by & ~1) + 2, bz & ~1:(bz & ~1) + 2, bx & ~1:(bx & ~1) + 2]
                    edge[sl] = 32 if not alloc.any() else (16 if not g.any() else 8)
                    continue
                desc[sl] = True
                for y in range(by * 8, by * 8 + 8):
                    for z in range(bz * 8, bz * 8 + 8):
                        for x in range(bx * 8, bx * 8 + 8):
                            if occ[y, z, x]:
                                solid[y, z, x] = True
                                continue
                            c4 = occ[y & ~3:(y & ~3) + 4, z & ~3:(z & ~3) + 4, x & ~3:(x & ~3) + 4]
                            c2 = occ[y & ~1:(y & ~1) + 2, z & ~1:(z & ~1) + 2, x & ~1:(x & ~1) + 2]
                            edge[y, z, x] = 4 if not c4.any() else (2 if not c2.any() else 1)
    return solid, desc, edge


# one direction per octant, and octants with a signed zero (d = -0.0 is not < 0)
DIRECTIONS = [(sx * 0.3, sy * 0.9, sz * 0.5) for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)] + \
             [(-0.0, -0.7, 0.7), (0.6, -0.0, -0.8), (-0.5, 0.5, -0.0)]


def all_positions():
    # rows ordered [y][z][x] to match the dense grid's flattening
    yy, zz, xx = np.meshgrid(np.arange(32), np.arange(32), np.arange(32), indexing="ij")
    return np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1).astype(np.int32)


def mismatches(step_fn, mask: int, vox: np.ndarray) -> int:
    """Steps of the oracle (step_fn(v, d) -> v_out, solid, descended) that differ from the
    closed form, over every voxel of the sector and every direction."""
    solid, desc, edge = expected_steps(mask, vox)
    v = all_positions()
    solid_f, desc_f, edge_f = solid.ravel(), desc.ravel(), edge.ravel()
    bad = 0
    for d in DIRECTIONS:
        dv = np.tile(np.array(d, np.float32), (v.shape[0], 1))
        out, s, de = step_fn(v, dv)
        exp = v.copy()
        for a in range(3):
            m = edge_f - 1
            lo, hi = v[:, a] & ~m, v[:, a] | m
            exp[:, a] = np.where(solid_f, v[:, a], lo if d[a] < 0 else hi)
        bad += int(np.count_nonzero(s != solid_f) + np.count_nonzero(de != desc_f) +
                   np.count_nonzero(np.any(out != exp, axis=1)))
    return bad


def sector_world(world_cls, mask: int, vox: np.ndarray):
    bricks = [vox[brick_slice(b)].reshape(-1) for b in range(64) if (mask >> b) & 1]
    br = np.concatenate(bricks) if bricks else np.zeros(0, np.uint8)
    return world_cls(5, 5, np.array([0], np.uint32), np.array([mask], np.uint64), br)


# ---------------------------------------------------------------- the pin


@pytest.mark.parametrize("kind,seed", SCENARIOS)
def test_step_pos_matches_closed_form_exhaustively(oracle_mod, kind, seed):
    mask, vox = scenario(kind, seed)
    ow = sector_world(oracle_mod.OracleWorld, mask, vox)
    assert mismatches(ow.step_pos, mask, vox) == 0


# ---------------------------------------------------------------- planted mistakes


class _MutantWorld:
    """A one-sector world of a separately compiled oracle library (the planted-mistake copy)."""

    def __init__(self, h, log2_xz, log2_y, si, am, br):
        self.h = h
        out = C.c_void_p()
        rc = h.oracle_world_create(log2_xz, log2_y, len(si), si.ctypes.data, am.ctypes.data,
                                   br.ctypes.data if br.size else None, C.byref(out))
        assert rc == 0
        self._w = out

    def step_pos(self, v, d):
        v = np.ascontiguousarray(v, np.int32)
        d = np.ascontiguousarray(d, np.float32)
        n = v.shape[0]
        out = np.zeros((n, 3), np.int32)
        s = np.zeros(n, np.uint8)
        de = np.zeros(n, np.uint8)
        self.h.oracle_step_pos(self._w, n, v.ctypes.data, d.ctypes.data, out.ctypes.data, s.ctypes.data,
                               de.ctypes.data)
        return out, s.astype(bool), de.astype(bool)

    def __del__(self):
        self.h.oracle_world_destroy(self._w)


# (text in traverse.c, planted replacement): the 2x2x2 group shift mask, the "+2 when the
# 64-bit mask is 0" tested on the 32-bit half, the +1 group level dropped (a too
# conservative LOD: hits unchanged, more iterations), and d = -0.0 taken as negative
MUTATIONS = [
    ("half >> (bi & 0xA)", "half >> (bi & 0x5)"),
    ("if (m == 0)", "if (half == 0)"),
    ("        lod += 1;", "        lod += 0;"),
    ("v[a] = d[a] < 0 ?", "v[a] = signbit(d[a]) ?"),
]


def build_mutant(tmp: Path, old: str, new: str) -> C.CDLL:
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    for f in ORACLE_DIR.iterdir():
        if f.suffix in (".c", ".h", ".inc"):
            shutil.copy(f, tmp / f.name)
    src = (tmp / "traverse.c").read_text()
    assert src.count(old) == 1, old
    (tmp / "traverse.c").write_text(src.replace(old, new))
    lib = tmp / "libmutant.so"
    srcs = [str(tmp / n) for n in ("traverse.c", "extras.c", "simd.c", "shader.c", "render.c")]
    subprocess.run([gcc, "-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-march=x86-64",
                    "-shared", "-o", str(lib), *srcs, "-lm", "-lpthread"], check=True, capture_output=True)
    h = C.CDLL(str(lib))
    P, u32 = C.c_void_p, C.c_uint32
    h.oracle_world_create.argtypes = [u32, u32, u32, P, P, P, C.POINTER(P)]
    h.oracle_world_create.restype = C.c_int
    h.oracle_world_destroy.argtypes = [P]
    h.oracle_step_pos.argtypes = [P, u32, P, P, P, P, P]
    return h


@pytest.mark.parametrize("old,new", MUTATIONS)
def test_planted_mistake_fails_the_pin(tmp_path, old, new):
    h = build_mutant(tmp_path, old, new)
    total = 0
    for kind, seed in SCENARIOS:
        mask, vox = scenario(kind, seed)
        w = sector_world(lambda *a: _MutantWorld(h, *a), mask, vox)
        total += mismatches(w.step_pos, mask, vox)
    assert total > 0, f"planted mistake {new!r} was not detected"


# ---------------------------------------------------------------- hand-counted iterations


def test_cpu_method_iteration_sequence_hand_counted(oracle_mod):
    """One solid voxel at (45, 1, 1) of a 64^3 world; nearly axis-aligned rays along x.

    +x from (0.5, 1.5, 1.5), lookups by hand:
      1. (0,1,1)  sector 0 has no brick: mask 0, LOD 3+2, 32-voxel cell -> x 31, t 31.5
      2. (32,1,1) brick 0 of sector 1 empty, its 2x2x2 brick group holds brick 1: LOD 3 -> x 39
      3. (40,1,1) brick 1 allocated (descent), cell x 40..43 empty: LOD 0+2 -> x 43
      4. (44,1,1) descent, cell x 44..47 holds the voxel, group {44,45}x{0,1}x{0,1} too: LOD 0
      5. (45,1,1) descent, solid: hit after 5 lookups, 3 descents, t = 44.5 (side x of step 4)
    -x from (63.5, 1.5, 1.5):
      1. (63,1,1) brick 3 empty, group {2,3}x{0,1}x{0,1} of bricks empty: LOD 3+1 -> x 48, t 15.5
      2. (47,1,1) descent, cell x 44..47 not empty, group {46,47}x{0,1}x{0,1} empty: LOD 1 -> x 46
      3. (45,1,1) descent, solid: 3 lookups, 2 descents, t = 17.5
    """
    from test_oracle import dense_world

    vox = np.zeros((64, 64, 64), np.uint8)
    vox[1, 1, 45] = 3
    ow = dense_world(oracle_mod, vox)
    e = np.float32(2.0 ** -20)
    o = np.array([[0.5, 63.5], [1.5, 1.5], [1.5, 1.5]], np.float32)
    d = np.array([[1.0, -1.0], [e, e], [e, e]], np.float32)
    r = ow.cast(o, d, (0, 0, 0), 128, with_descents=True)
    rec = r["rec"]
    assert tuple(rec[0, :3]) == (45, 1, 1) and tuple(rec[1, :3]) == (45, 1, 1)
    assert [int(x) >> 16 for x in rec[:, 3]] == [5, 3]
    assert list(r["descents"]) == [3, 2]
    assert list(r["t"]) == [44.5, 17.5]
    assert [(int(x) >> 8) & 0x3F for x in rec[:, 3]] == [3, 1]  # normal x: -1 going +x, +1 going -x
    c = ow.count(o, d, (0, 0, 0), 128)
    assert (c["iters"], c["descents"], c["hits"]) == (8, 5, 2)


# ---------------------------------------------------------------- PCG known answers


def test_pcg_known_answers(oracle_mod):
    """random_u32 (RandomGen.glsl:6-11) returns the output permutation of the state before the
    LCG step: word = ((s >> ((s >> 28) + 4)) ^ s) * 277803737, out = (word >> 22) ^ word.
    Values worked by hand:
      s = 0:          shift 4, word 0, out 0; next state = the increment 2891336453
      s = 1:          shift 4, 1 >> 4 = 0, word = 277803737 = 0x108EF2D9, word >> 22 = 0x42,
                      out = 0x108EF2D9 ^ 0x42 = 0x108EF29B
      s = 0x10000000: shift 5, s ^ (s >> 5) = 0x10800000 = 2^28 + 2^23, so word =
                      (0x108EF2D9 << 28) + (0x108EF2D9 << 23) mod 2^32
                      = 0x90000000 + (0x0D9 << 23 = 0x6C800000) = 0xFC800000; out = 0xFC800000 ^ 0x3F2
    random_init(uvec3 p) (:16-20): n = (1103515245 p.x) ^ ((1103515245 p.y) >> 3), + p.z * 12345:
      (1, 0, 0) -> 1103515245 = 0x41C64E6D; (0, 8, 0) -> (8 * 0x41C64E6D mod 2^32) >> 3 = 0x01C64E6D
      (0x41C64E6D with its top 3 bits dropped); (0, 0, 1) -> 12345; (0, 0, 2) -> 24690."""
    assert oracle_mod.pcg_next(0) == (0, 2891336453)
    assert oracle_mod.pcg_next(1)[0] == 0x108EF29B
    assert oracle_mod.pcg_next(1)[1] == 747796405 + 2891336453  # 3639132858 < 2^32: no wrap
    assert oracle_mod.pcg_next(0x10000000)[0] == 0xFC8003F2
    assert oracle_mod.pcg_init(1, 0, 0) == 0x41C64E6D
    assert oracle_mod.pcg_init(0, 8, 0) == 0x01C64E6D
    assert oracle_mod.pcg_init(0, 0, 1) == 12345
    assert oracle_mod.pcg_init(0, 0, 2) == 24690
    assert oracle_mod.pcg_init(1, 8, 1) == (0x41C64E6D ^ 0x01C64E6D) + 12345
