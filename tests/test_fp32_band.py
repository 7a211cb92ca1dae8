"""The FP32 filter of the Poisson distance test and its error bound, on the CPU.

The kernels decide `dd < spacing^2` (sampler_poisson.h:160-166) in FP32 and
re-check in FP64 only pairs whose FP32 value lies in the band
[0.999, 1.001] * spacing^2.  The kernels are compiled without FMA
contraction, so their FP32 arithmetic is IEEE round-to-nearest + - *, which
numpy float32 reproduces operation for operation.  This file restates the
two FP32 expressions (k_ring: coordinates relative to the node centre;
k_neighbors / k_accept / k_sequential: integer deltas times the FP32 scale),
checks the error bound that DESIGN.md derives against brute force at the
band edges, and checks that no pair outside the band is decided differently
from the reference's FP64 expression, at the coordinate ranges of the
workloads (UTM-like offsets, the precision scene's 1e6 m offset at scale
0.0001) and down to spacings of 100 scale units.

Bound (u = 2^-24, h = 64 spacings, the half size of a node, D the exact
squared distance, D <= 1.001 s^2):
  centre-relative:  |d32 - D| <= ((2 * (2h/s + 1) * sqrt(3) + 3) u) s^2
                                <= 450 u s^2 ~ 2.7e-5 s^2
  integer deltas:   |d32 - D| <= 9 u D ~ 5.4e-7 s^2
and the FP64 reference differs from D by at most 4 sqrt(3) eps |x|max
sqrt(D) (eps = 2^-53, coordinates rounded once each by X * scale + offset),
far below the band's 1e-3 s^2 half width whenever s > 1e-9 |x|max.
"""
import numpy as np
import pytest

U = 2.0 ** -24
BOUND_RING = 450 * U
BOUND_DELTA = 9 * U
F32 = np.float32


def d64(a, b, scale, offset):
    """The reference expression: (X * scale + offset) per point, then the
    squared differences in double (sampler_poisson.h:155-159)."""
    xa = a * scale + offset
    xb = b * scale + offset
    d = xa - xb
    return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]


def d32_ring(a, b, scale, offset, centre):
    """k_ring: float(coordinate - centre) per point, then FP32 differences."""
    fa = ((a * scale + offset) - centre).astype(F32)
    fb = ((b * scale + offset) - centre).astype(F32)
    d = fa - fb
    return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]


def d32_delta(a, b, scale):
    """within_xyz: float(int64 delta) * float(scale)."""
    sc = np.asarray(scale).astype(F32)
    d = (a.astype(np.int64) - b.astype(np.int64)).astype(F32) * sc
    return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]


def pairs(rng, n, s, scale, offset, centre_q, h):
    """Point a anywhere in the node (|x - c| <= h, plus one unit), b at
    distance 0.995 .. 1.005 s from a in a random direction, both quantized
    to the LAS grid."""
    a = centre_q + rng.uniform(-h, h, size=(n, 3)) / scale
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    r = s * rng.uniform(0.995, 1.005, size=(n, 1))
    b = a + v * r / scale
    return np.rint(a), np.rint(b)


CASES = [
    # scale, offset, spacing: UTM tile root and deep levels, the precision
    # scene (1e6 m at 0.0001), and a spacing of 100 LAS units
    (0.001, 5_000_000.0, 39.0625),
    (0.001, 5_000_000.0, 39.0625 / 2 ** 10),
    (0.0001, 1_000_000.0, 7.8125),
    (0.0001, 1_000_000.0, 7.8125 / 2 ** 12),
    (0.0001, 1_000_000.0, 0.01),
    (0.01, 0.0, 1.0),
]


@pytest.mark.parametrize("scale,offset,s", CASES, ids=[f"sc{c[0]}_off{c[1]:.0f}_s{c[2]:.4g}" for c in CASES])
def test_band_decisions_equal_the_fp64_decisions(scale, offset, s):
    rng = np.random.default_rng(17)
    n = 400_000
    h = 64 * s
    centre = offset + 1000.0 + rng.uniform(0, 1000.0)  # node centre in world coordinates
    centre_q = (centre - offset) / scale
    a, b = pairs(rng, n, s, scale, offset, centre_q, h)
    s2 = s * s
    ref = d64(a, b, scale, offset) < s2
    s2f = F32(s2)
    lo, hi = s2f * F32(0.999), s2f * F32(1.001)
    for d32, bound in ((d32_ring(a, b, scale, offset, centre), BOUND_RING), (d32_delta(a, b, scale), BOUND_DELTA)):
        # the error bound holds
        D = d64(a, b, scale, offset)
        err = np.abs(d32.astype(np.float64) - D) / s2
        assert err.max() <= bound, (err.max(), bound)
        # outside the band the FP32 decision is the FP64 one
        inside = d32 < lo
        outside = d32 > hi
        assert np.array_equal(ref[inside], np.ones(inside.sum(), bool))
        assert np.array_equal(ref[outside], np.zeros(outside.sum(), bool))
        # the band is entered at all (the sample straddles it)
        assert (~inside & ~outside).sum() > 0


def test_bound_is_reached_in_order_of_magnitude():
    """The centre-relative bound is not loose by orders of magnitude: pairs
    at the node's corner (|x - c| = h on every axis) reach a third of it."""
    rng = np.random.default_rng(5)
    scale, offset, s = 0.0001, 1_000_000.0, 7.8125 / 2 ** 6
    h = 64 * s
    centre = offset + 1234.5678
    n = 200_000
    a = (centre - offset + np.sign(rng.uniform(-1, 1, size=(n, 3))) * h * 0.999) / scale
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    b = a + v * s / scale
    a, b = np.rint(a), np.rint(b)
    err = np.abs(d32_ring(a, b, scale, offset, centre).astype(np.float64) - d64(a, b, scale, offset)) / (s * s)
    assert BOUND_RING / 20 < err.max() <= BOUND_RING


def test_band_edges_themselves():
    """A value one FP32 ulp beyond either band edge decides like FP64 for
    every pair whose FP32 value lands there."""
    rng = np.random.default_rng(9)
    scale, offset, s = 0.0001, 1_000_000.0, 0.05
    centre = offset + 500.0
    a, b = pairs(rng, 2_000_000, s * 1.0, scale, offset, (centre - offset) / scale, 64 * s)
    s2 = s * s
    s2f = F32(s2)
    lo, hi = s2f * F32(0.999), s2f * F32(1.001)
    d32 = d32_ring(a, b, scale, offset, centre)
    ref = d64(a, b, scale, offset) < s2
    near_lo = (d32 < lo) & (d32 >= np.nextafter(lo, F32(0), dtype=F32) * F32(0.9999))
    near_hi = (d32 > hi) & (d32 <= np.nextafter(hi, F32(1e30), dtype=F32) * F32(1.0001))
    assert near_lo.sum() > 0 and near_hi.sum() > 0
    assert ref[near_lo].all() and not ref[near_hi].any()
