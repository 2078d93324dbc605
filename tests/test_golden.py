"""The oracle reproduces its committed golden fixtures exactly (tests/golden/make_golden.py wrote them), and the
fixtures hold the properties the path guarantees: equal-seed fp32/fp64 agreement on primary hits."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import CASES, make

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(name):
    g = np.load(os.path.join(HERE, f"{name}.npz"))
    now = make(name)
    for k in g.files:
        assert np.array_equal(g[k], now[k]), k


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_primary_hits_agree_across_precisions(name):
    """Equal-seed primary hits of the fp32 path against the fp64 engine precision.  Meshes and SDFs meet 1e-5
    relative on distance at the 99th percentile.  Sphere scenes do not: the engine's c = |oc|^2 - r^2
    (sphere.rs:151) cancels in fp32 for the r = 1000 ground sphere, leaving up to ~3e-4."""
    g = np.load(os.path.join(HERE, f"{name}.npz"))
    same = g["rank"] == g["rank64"]
    assert same.mean() > 0.99
    hit = same & (g["rank"] != -1)
    rel = np.abs(g["t32"][hit] - g["t64"][hit]) / np.abs(g["t64"][hit])
    if name in ("random_spheres", "mixed_primitives"):
        assert np.median(rel) <= 1e-5 and rel.max() <= 1e-3
    else:
        assert np.quantile(rel, 0.99) <= 1e-5 and rel.max() <= 1e-4
