"""Pins the oracle's arithmetic building blocks: the counter-based RNG and the single-precision functions
(oracle/twin_math.hpp, oracle/twin_rng.hpp) that the GPU kernels restate operation for operation.

Philox is pinned three ways: the published Random123 known-answer vectors for Philox4x32-10, an independent
Python implementation of the round structure validated against numpy's Philox4x64-10, and that same Python
structure at 32 bits against the C++ twin.  The math functions are pinned against the C library in double
precision with explicit error bounds and exact special values."""
import math

import numpy as np
import pytest

import oracle


def philox_py(ctr, key, w, rounds=10):
    if w == 64:
        m0, m1, w0, w1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157, 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
    else:
        m0, m1, w0, w1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    mask = (1 << w) - 1
    c, k = list(ctr), list(key)
    for r in range(rounds):
        if r:
            k = [(k[0] + w0) & mask, (k[1] + w1) & mask]
        p0, p1 = m0 * c[0], m1 * c[2]
        c = [(p1 >> w) ^ c[1] ^ k[0], p1 & mask, (p0 >> w) ^ c[3] ^ k[1], p0 & mask]
    return c


# Random123 kat_vectors, philox4x32 10 rounds: (counter, key, output)
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox4x32_known_answers(ctr, key, out):
    assert oracle.philox4x32(ctr, key) == out
    assert philox_py(ctr, key, 32) == out


def test_philox_structure_matches_numpy_4x64():
    key = [0x0123456789ABCDEF, 0xFEDCBA9876543210]
    for c0 in (0, 1, 41, 2**40 + 3):
        bg = np.random.Philox(key=np.array(key, dtype=np.uint64), counter=np.array([c0, 7, 0, 9], dtype=np.uint64))
        # numpy increments the counter before each block
        assert [int(x) for x in bg.random_raw(4)] == philox_py([c0 + 1, 7, 0, 9], key, 64)


def test_philox4x32_twin_matches_python_structure():
    rng = np.random.default_rng(3)
    for _ in range(200):
        ctr = [int(x) for x in rng.integers(0, 2**32, 4)]
        key = [int(x) for x in rng.integers(0, 2**32, 2)]
        assert oracle.philox4x32(ctr, key) == philox_py(ctr, key, 32)


def _ulp_err(got, want):
    got = np.asarray(got, np.float64)
    ulp = np.spacing(np.abs(np.asarray(want, np.float32)).astype(np.float32)).astype(np.float64)
    return np.abs(got - want) / np.maximum(ulp, np.finfo(np.float32).tiny)


def test_sincos_turns():
    r = np.concatenate([np.linspace(0, 1, 200001, endpoint=False), [0.125, 0.25, 0.375, 0.5, 0.75]]).astype(np.float32)
    s = oracle.math_f32("sin_turns", r)
    c = oracle.math_f32("cos_turns", r)
    a = 2 * np.pi * r.astype(np.float64)
    assert np.max(np.abs(s - np.sin(a))) < 2.5e-7
    assert np.max(np.abs(c - np.cos(a))) < 2.5e-7
    # exact quadrant values
    s4 = oracle.math_f32("sin_turns", np.array([0.0, 0.25, 0.5, 0.75], np.float32))
    c4 = oracle.math_f32("cos_turns", np.array([0.0, 0.25, 0.5, 0.75], np.float32))
    assert list(s4) == [0.0, 1.0, 0.0, -1.0]
    assert list(c4) == [1.0, 0.0, -1.0, 0.0]


def test_log():
    x = np.concatenate([np.arange(1, 1 << 16) * 2.0**-24, np.geomspace(1e-30, 1e30, 100001)]).astype(np.float32)
    got = oracle.math_f32("log", x)
    assert np.max(_ulp_err(got, np.log(x.astype(np.float64)))) <= 2.0
    sp = oracle.math_f32("log", np.array([1.0, 0.0, -1.0, np.inf, 1e-40], np.float32))
    assert sp[0] == 0.0 and sp[1] == -np.inf and np.isnan(sp[2]) and sp[3] == np.inf
    assert abs(sp[4] - math.log(np.float32(1e-40))) < 1e-4


def test_exp():
    x = np.linspace(-87, 88, 400001).astype(np.float32)
    got = oracle.math_f32("exp", x)
    assert np.max(_ulp_err(got, np.exp(x.astype(np.float64)))) <= 2.0
    sp = oracle.math_f32("exp", np.array([0.0, -0.0, 100.0, -200.0], np.float32))
    assert list(sp) == [1.0, 1.0, np.inf, 0.0]


def test_cbrt01():
    x = np.concatenate([np.arange(1, 1 << 16) * 2.0**-16, np.geomspace(2.0**-24, 1, 50001)]).astype(np.float32)
    got = oracle.math_f32("cbrt01", x)
    assert np.max(_ulp_err(got, np.cbrt(x.astype(np.float64)))) <= 2.0
    cubes = (np.arange(1, 65) / 64.0) ** 3
    assert np.array_equal(oracle.math_f32("cbrt01", cubes.astype(np.float32)), (np.arange(1, 65) / 64.0).astype(np.float32))
    assert oracle.math_f32("cbrt01", np.array([0.0], np.float32))[0] == 0.0


def test_sincos_radians():
    x = np.linspace(-9 * np.pi, 9 * np.pi, 400001).astype(np.float32)
    xd = x.astype(np.float64)
    assert np.max(np.abs(oracle.math_f32("sin", x) - np.sin(xd))) < 4e-7
    assert np.max(np.abs(oracle.math_f32("cos", x) - np.cos(xd))) < 4e-7


def test_atan2_and_acos():
    rng = np.random.default_rng(1)
    y = rng.normal(size=200000).astype(np.float32)
    x = rng.normal(size=200000).astype(np.float32)
    got = oracle.math_f32("atan2", y, x)
    assert np.max(np.abs(got - np.arctan2(y.astype(np.float64), x.astype(np.float64)))) < 4e-7
    # quadrant and signed-zero cases
    cases = [(0.0, 1.0, 0.0), (0.0, -1.0, np.float32(np.pi)), (1.0, 0.0, np.float32(np.pi / 2)),
             (-1.0, 0.0, -np.float32(np.pi / 2)), (-0.0, -1.0, -np.float32(np.pi))]
    for yy, xx, want in cases:
        assert oracle.math_f32("atan2", np.array([yy], np.float32), np.array([xx], np.float32))[0] == want
    c = np.linspace(-1, 1, 200001).astype(np.float32)
    assert np.max(np.abs(oracle.math_f32("acos", c) - np.arccos(c.astype(np.float64)))) < 6e-7
    assert oracle.math_f32("acos", np.array([1.0], np.float32))[0] == 0.0
