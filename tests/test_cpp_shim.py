"""The C++ shim (include/vrt.hpp) over the C ABI, compiled with g++ as the reference's own C++
callers would: the error path on the CPU, the picking cast (VoxelMap::RayCast) on the GPU."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def build_shim(tmp_path: Path) -> Path:
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "shim_check"
    lib_dir = ROOT / "voxelrt_amd"
    cmd = [cxx, "-std=c++17", "-O1", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(ROOT / "tests/cpp/shim_check.cpp"),
           f"-L{lib_dir}", "-lvoxelrt_hip", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


def test_shim_reports_missing_device(vrt, tmp_path):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    exe = build_shim(tmp_path)
    res = subprocess.run([str(exe), "nodevice"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "status 4" in res.stdout


@pytest.mark.gpu
def test_shim_picking_cast_matches_oracle(vrt, oracle_mod, cuda, tmp_path):
    from helpers import blob_world, random_rays
    from test_oracle import dense_world

    vox = blob_world(64, seed=17, density=0.004, blobs=3)
    ow = dense_world(oracle_mod, vox)
    o, d = random_rays(200, 64, seed=18, inside=False)
    rays = np.concatenate([o, d]).T.astype(np.float64)
    (tmp_path / "vox.bin").write_bytes(np.ascontiguousarray(vox).tobytes())
    (tmp_path / "rays.bin").write_bytes(np.ascontiguousarray(rays).tobytes())
    exe = build_shim(tmp_path)
    res = subprocess.run([str(exe), "pick", str(tmp_path / "vox.bin"), "64", str(tmp_path / "rays.bin")],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    lines = res.stdout.strip().split("\n")
    assert len(lines) == 200
    hits = 0
    for line, ray in zip(lines, rays):
        f = line.split()
        ref = ow.cast_f64(ray[:3], ray[3:], 1024)
        assert float(f[0]) == ref["distance"]
        assert tuple(int(x) for x in f[1:4]) == ref["voxel"]
        assert tuple(np.float32(x) for x in f[4:7]) == tuple(np.float32(x) for x in ref["normal"])
        assert tuple(np.float32(x) for x in f[7:9]) == tuple(np.float32(x) for x in ref["uv"])
        hits += ref["distance"] > 0
    assert hits > 20
