"""Seeded synthetic workloads: worlds, cameras and ray sets.

    C1  256^3 fBm terrain (4 octaves, frequency 1/64, gain 0.5), 640x360 primary rays
    C2  1024^3 fBm terrain + 3D value-noise caves (solid where noise > 0.55),
        1920x1080 primary rays + one diffuse bounce per hit
    C3  2048^3, 20k spheres r ~ U(4, 64) (the literal set, ~49% fill) or a
        sparse set of 587 spheres (~2% fill); 1920x1080 primary + 2M incoherent rays
    C4  the C2 world, 64 cameras of uneven square sizes + bounce sets (batched)
    C5  512^3 terrain, 16 views of 512x512, per-voxel RGB albedo (differentiable)

The C2, C4 and C5 cameras put their rays' integer world origin at the corner of the
camera's sector (a multiple of 32 per axis: the casts' aligned-origin step); C1 and C3
at floor(eye), the reference's choice (the general step).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .api import CameraSpec, Scene

MAX_ITERS = 128          # the reference CPU cast's iteration cap (CpuRenderer.cpp:185)
BOUNCE_SEED = 1          # PCG frame seed of the bounce rays


@dataclass(frozen=True)
class TerrainSpec:
    log2_xz: int
    log2_y: int
    seed: int
    octaves: int = 4
    frequency: float = 1 / 64
    gain: float = 0.5
    base: float = 32.0
    amp: float = 128.0
    cave_threshold: float = 0.0
    cave_octaves: int = 3
    cave_frequency: float = 1 / 32
    band: int = 16

    def build(self, threads: int = 0) -> Scene:
        return Scene.terrain(self.log2_xz, self.log2_y, self.seed, octaves=self.octaves, frequency=self.frequency,
                             gain=self.gain, base=self.base, amp=self.amp, cave_threshold=self.cave_threshold,
                             cave_octaves=self.cave_octaves, cave_frequency=self.cave_frequency, band=self.band,
                             threads=threads)


# C1: 256^3, heightmap between 32 and 160 voxels.
C1_WORLD = TerrainSpec(8, 8, seed=1, base=32.0, amp=128.0)
# C2: 1024^3, heightmap 32..332 voxels carved by caves; ~12% of 128^3 bricks allocated.
C2_WORLD = TerrainSpec(10, 10, seed=2, base=32.0, amp=300.0, frequency=1 / 256, cave_threshold=0.55,
                       cave_octaves=3, cave_frequency=1 / 32)
# C5: the C1 recipe scaled to 512^3.
C5_WORLD = TerrainSpec(9, 9, seed=5, base=64.0, amp=256.0, frequency=1 / 128)


def c1_camera(tile=(8, 8), width=640, height=360) -> CameraSpec:
    return CameraSpec((128.37, 200.61, 12.29), yaw_deg=8.0, pitch_deg=-28.0, width=width, height=height, tile=tile)


def c2_camera(tile=(8, 8), width=1920, height=1080, anchor="sector") -> CameraSpec:
    return CameraSpec((512.37, 420.61, 96.29), yaw_deg=12.0, pitch_deg=-24.0, width=width, height=height, tile=tile,
                      anchor=anchor)


def c3_world(sparse: bool, threads: int = 0) -> Scene:
    if sparse:
        return Scene.spheres(11, 11, 587, 4.0, 64.0, seed=3, threads=threads)
    return Scene.spheres(11, 11, 20000, 4.0, 64.0, seed=3, threads=threads)


def c3_camera(tile=(8, 8)) -> CameraSpec:
    """Inside the box near its z = 0 face, looking across it."""
    return CameraSpec((1024.37, 1100.61, 6.29), yaw_deg=3.0, pitch_deg=-4.0, width=1920, height=1080, tile=tile)


def c4_cameras(n: int = 64, seed: int = 4, tile=(8, 8)) -> list[CameraSpec]:
    """64 cameras, square sizes uniform over multiples of 16 in [128, 1024], seeded poses above the terrain."""
    rng = np.random.default_rng(seed)
    cams = []
    for _ in range(n):
        size = int(rng.integers(8, 65)) * 16
        pos = (float(rng.uniform(64, 960)), float(rng.uniform(350, 500)), float(rng.uniform(64, 960)))
        yaw = float(rng.uniform(0, 360))
        pitch = float(rng.uniform(-50, -10))
        cams.append(CameraSpec(pos, yaw, pitch, size, size, tile=tile, anchor="sector"))
    return cams


def c5_cameras(n: int = 16, size: int = 512, tile=(8, 8), anchor: str = "sector") -> list[CameraSpec]:
    cams = []
    for k in range(n):
        a = 2 * math.pi * k / n
        pos = (256 + 200 * math.sin(a) + 0.31, 420.57, 256 + 200 * math.cos(a) + 0.17)
        yaw = math.degrees(a) + 180.0
        cams.append(CameraSpec(pos, yaw, -35.0, size, size, tile=tile, anchor=anchor))
    return cams


def brick_fraction(scene: Scene) -> float:
    s = scene.stream()
    total = (1 << (s["log2_size_xz"] - 3)) ** 2 * (1 << (s["log2_size_y"] - 3))
    return s["nbricks"] / total
