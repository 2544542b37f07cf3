"""Writes tests/golden/blob64.npz: a seeded 64^3 world, 4096 rays and the oracle's hits.

The reference has no tests or golden vectors for this path (its traversal
needs glm/imgui/GL to build), so the fixture is produced by the oracle and
pins it, and the GPU path, to these exact bytes on every machine.
Run from the repository root: python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent / "oracle"))

from helpers import blob_world, random_rays  # noqa: E402
from test_oracle import dense_world  # noqa: E402
import oracle  # noqa: E402


def main():
    vox = blob_world(64, seed=21, density=0.01, blobs=5)
    W = np.array([5, -3, 7], np.int32)
    o, d = random_rays(4096, 64, seed=22, inside=False)
    o = (o - W[:, None]).astype(np.float32)
    ow = dense_world(oracle, vox)
    r = ow.cast(o, d, tuple(W), 128)
    np.savez_compressed(HERE / "blob64.npz", voxels=vox, o=o, d=d, world_origin=W, max_iters=np.int32(128),
                        t=r["t"], rec=r["rec"])
    hit = (r["rec"][:, 3] & 0xFF) != 0
    print(f"wrote blob64.npz: {hit.sum()} hits of {len(hit)}")


if __name__ == "__main__":
    main()
