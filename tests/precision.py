"""The equal-seed precision contract between the single-precision path (the GPU, bit for bit the oracle's fp32
restatement) and the engine's double precision (the oracle's fp64 restatement, engine boxes, recursive
integrator), at the same RNG counters:

* primary-hit primitives (depth-first ranks) agree on at least RANK_MIN of the pixels;
* primary-hit distances agree within T_REL relative, except at grazing incidence (|d.n| < GRAZING_COS), where
  the distance's condition number 1/|d.n| magnifies the ~6e-8 difference between the two precisions' camera
  rays;
* per-pixel radiance agrees within L_REL relative, except at pixels where some sample's path
  - diverges: the precisions' paths hit different primitives or end differently (a branch such as Schlick's
    reflectance against a uniform draw, or a closest hit decided by less than the rounding);
  - meets an SDF: its shading normal is a central difference over 2 epsilon = 2e-4 of the distance estimate,
    whose single-precision rounding (~1e-7 of the distance) leaves normals ~1e-3 apart;
  - reflects or refracts (metal, dielectric) and bounces on: mirror chains between curved surfaces magnify
    direction differences by ~2 d / r per bounce.
Each exception class is counted per workload, and the fractions are bounded by the callers.
"""
import numpy as np

import oracle

T_REL = 1e-5
L_REL = 1e-4
RANK_MIN = 0.999
GRAZING_COS = 0.1


def primary_check(tw, cam, opts, rank32, t32, rank64, t64):
    """Counts for the primary hits of sample 0: pixels, rank agreements, hits, distance exceptions (all at
    grazing incidence: asserted)."""
    same = rank32 == rank64
    hit = same & (rank32 != -1)
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = np.abs(t32.astype(np.float64) - t64) / np.abs(t64)
    off = hit & ~(rel <= T_REL)
    grazing = 0
    for y, x in zip(*np.nonzero(off)):
        o, d = oracle.camera_ray(cam, opts, int(x), int(y), 0)
        h = tw.trace(o, d)
        assert h is not None, (y, x)
        cos = abs(float(np.dot(d, h["normal"])))
        assert cos < GRAZING_COS, f"pixel {(y, x)}: distance off by {rel[y, x]:.2e} at |d.n| = {cos:.3f}"
        grazing += 1
    return dict(pixels=int(same.size), rank_same=float(same.mean()), hits=int(hit.sum()), t_grazing=grazing,
                t_max_rel=float(np.max(rel[hit & (rel <= T_REL)], initial=0.0)))


def radiance_check(tw, scene, cam, opts, img32, img64):
    """Counts of the pixels outside L_REL, by cause; every one must have a cause (asserted)."""
    rel = np.abs(img32.astype(np.float64) - img64) / np.maximum(np.abs(img64.astype(np.float64)), 1e-6)
    off = ~(rel <= L_REL).all(axis=2)
    obj, _ = tw.dfs_order()  # the object (an index into scene.objects) of each depth-first rank
    specular = {"MetalMaterial", "DielectricMaterial"}

    def kind(r):
        return type(scene.objects[obj[r]].material).__name__

    counts = dict(diverged=0, sdf=0, specular_chain=0)
    for y, x in zip(*np.nonzero(off)):
        p32 = tw.path_ranks(cam, opts, int(x), int(y), 32)
        p64 = tw.path_ranks(cam, opts, int(x), int(y), 64)
        if p32 != p64:
            counts["diverged"] += 1
            continue
        if any(r <= -2 and r > oracle.PATH_SKY for p in p32 for r in p):
            counts["sdf"] += 1
            continue
        if any(r >= 0 and kind(r) in specular and i + 1 < len(p) and p[i + 1] > oracle.PATH_SKY
               for p in p32 for i, r in enumerate(p)):
            counts["specular_chain"] += 1
            continue
        raise AssertionError(f"pixel {(y, x)}: radiance off by {rel[y, x].max():.2e} on identical paths "
                             f"without SDF or specular vertices: {p32}")
    counts["pixels"] = int(off.size)
    counts["off"] = int(off.sum())
    return counts
