"""The camera-ray packets' fused slab test is conservative.

k_extend_packet (rayna_amd/csrc/traverse.hpp) evaluates a box plane's slab distance as fma(b, 1/d, -o/d) and,
to cover the rounding of -o/d, lowers the entry distance and raises the exit distance by 2^-20 max |o/d|;
waves holding a ray with a zero direction component keep the two-rounding form (b - o) * (1/d).  Culling must
stay conservative: whenever the ray, in exact arithmetic, passes through the box at some t in [T_MIN, best],
the test must keep the box, or a closest hit could be lost.  This restates both single-precision tests -- the
fused multiply-add exactly, through 80-bit intermediates (the 48-bit product is exact there and the sum rounds
once to 64 bits before the rounding to 32) -- and checks them against the exact intervals on rays and boxes built
to stress the rounding: origins far from the scene's origin, boxes just touching the ray, planes within an ulp
of the origin, nearly parallel directions.  The GPU parity tests check the kernel itself against the oracle."""
import numpy as np

F = np.float32
T_MIN = F(1e-3)
SLACK = F(1.0 + 2.0 ** -16)
L = np.longdouble


def fma32(a, b, c):
    return (a.astype(L) * b.astype(L) + c.astype(L)).astype(F)


def two_rounding(o, inv, lo, hi, best):
    with np.errstate(all="ignore"):
        t1 = (lo - o) * inv
        t2 = (hi - o) * inv
        t0 = np.fmax.reduce(np.fmin(t1, t2), axis=1)
        tf = np.fmin.reduce(np.fmax(t1, t2), axis=1)
        return np.fmax(t0, T_MIN) <= np.fmin(tf, best) * SLACK


def exact(o, d, lo, hi, best):
    """The ray meets the box at some t in [T_MIN, best], in 80-bit arithmetic (d has no zero component here)."""
    oo, dd = o.astype(L), d.astype(L)
    t1 = (lo.astype(L) - oo) / dd
    t2 = (hi.astype(L) - oo) / dd
    t0 = np.max(np.minimum(t1, t2), axis=1)
    tf = np.min(np.maximum(t1, t2), axis=1)
    return np.maximum(t0, L(T_MIN)) <= np.minimum(tf, best.astype(L))


def fused(o, inv, lo, hi, best):
    with np.errstate(all="ignore"):
        noi = -(o * inv)
        slack = np.max(np.abs(noi), axis=1) * F(2.0 ** -20)
        t1 = fma32(lo, inv, noi)
        t2 = fma32(hi, inv, noi)
        t0 = np.fmax.reduce(np.fmin(t1, t2), axis=1) - slack
        tf = np.fmin.reduce(np.fmax(t1, t2), axis=1) + slack
        return np.fmax(t0, T_MIN) <= np.fmin(tf, best) * SLACK


def rays(rng, n, scale):
    o = (rng.uniform(-1, 1, (n, 3)) * scale).astype(F)
    d = rng.standard_normal((n, 3))
    tiny = rng.random((n, 3)) < 0.1  # nearly parallel to an axis
    d[tiny] *= 1e-6
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(F)
    return o, d, (F(1.0) / d).astype(F)


def test_fused_keeps_every_box_the_ray_meets():
    rng = np.random.Generator(np.random.PCG64(5))
    kept = 0
    for scale in (1e-3, 1.0, 50.0, 3e4):
        n = 200_000
        o, d, inv = rays(rng, n, scale)
        # boxes around a point on the ray at distance t, of random size, shifted so a face touches the ray
        t = (rng.uniform(0.0, 3.0, n) * max(scale, 1.0)).astype(F)
        p = (o.astype(np.float64) + d.astype(np.float64) * t[:, None])
        half = rng.uniform(0.0, 0.05, (n, 3)) * max(scale, 1.0)
        shift = rng.choice([-1.0, 0.0, 1.0], (n, 3)) * half * rng.uniform(0.99, 1.01, (n, 3))
        lo = np.nextafter((p - half + shift).astype(F), F(-np.inf))
        hi = np.nextafter((p + half + shift).astype(F), F(np.inf))
        # some planes within an ulp of the origin, some flat boxes
        on = rng.random((n, 3)) < 0.05
        lo = np.where(on, o, lo)
        flat = rng.random((n, 3)) < 0.05
        hi = np.where(flat, lo, hi)
        best = np.where(rng.random(n) < 0.5, F(np.finfo(F).max), (t * F(1.0001)).astype(F))
        e = exact(o, d, lo, hi, best)
        for test in (two_rounding, fused):
            lost = e & ~test(o, inv, lo, hi, best)
            assert not lost.any(), (test.__name__, scale, o[lost][:2], d[lost][:2], lo[lost][:2], hi[lost][:2])
        kept += int(e.sum())
    assert kept > 100_000


def test_fused_test_without_slack_would_lose_boxes():
    """The slack is needed: without it the rounding of -o/d loses boxes the ray meets."""
    rng = np.random.Generator(np.random.PCG64(6))
    n = 400_000
    o, d, inv = rays(rng, n, 3e4)
    t = rng.uniform(0.0, 1.0, n).astype(F)
    p = o.astype(np.float64) + d.astype(np.float64) * t[:, None]
    lo = (p - 1e-3).astype(F)
    hi = p.astype(F)
    lo = np.where(rng.random((n, 3)) < 0.5, hi, lo)  # flat boxes through the ray's point
    best = np.full(n, np.finfo(F).max, F)
    with np.errstate(all="ignore"):
        noi = -(o * inv)
        t1, t2 = fma32(lo, inv, noi), fma32(hi, inv, noi)
        t0 = np.fmax.reduce(np.fmin(t1, t2), axis=1)
        tf = np.fmin.reduce(np.fmax(t1, t2), axis=1)
        bare = np.fmax(t0, T_MIN) <= np.fmin(tf, best) * SLACK
    e = exact(o, d, lo, hi, best)
    assert (e & ~bare).any()
    assert not (e & ~fused(o, inv, lo, hi, best)).any()
