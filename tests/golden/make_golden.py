"""Writes the golden fixtures of tests/golden/ from the oracle (the CPU restatement of the render path).

    python tests/golden/make_golden.py

Each workload is rendered small by both twins: fp32 (the GPU's arithmetic, forward integrator, tight boxes) and
fp64 (the engine's precision and recursive integrator, engine boxes).  Stored per workload: primary-hit ranks
and distances of sample 0, both images, segment counts, and the BVH topology hash.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from rayna_amd import presets  # noqa: E402

CASES = {
    "random_spheres": lambda: presets.random_spheres(width=64, height=36, samples=4),
    "mixed_primitives": lambda: presets.mixed_primitives(width=48, height=32, samples=4),
    "triangle_mesh_s4": lambda: presets.triangle_mesh(subdiv=4, width=48, height=27, samples=2),
    "sdf_volume": lambda: presets.sdf_volume(width=32, height=18, samples=2),
}


def make(name):
    sc, cam, opts = CASES[name]()
    tw = oracle.TwinScene(sc)
    rank, t32 = tw.primary_hits(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    rank64, t64 = tw.primary_hits(cam, opts, precision=64, box_mode=oracle.BOX_REFERENCE)
    img32, segs32, _, _ = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    img64, segs64, _, _ = tw.render(cam, opts, precision=64, box_mode=oracle.BOX_REFERENCE)
    return dict(rank=rank, t32=t32.astype(np.float32), rank64=rank64, t64=t64, img32=img32, img64=img64,
                segs32=np.int64(segs32), segs64=np.int64(segs64), tree_hash=np.uint64(tw.tree_hash()))


def main():
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **make(name))
        print("wrote", name)


if __name__ == "__main__":
    main()
