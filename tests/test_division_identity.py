"""The identity the bit-exact walks use for Gravity::particleParticle's three
divisions by one denominator (n-body_amd/csrc/device.hpp, ieee_div3): for
FP32 n and d > 0 (finite),

    fl32(n / d) = fl32(fl64(n * r))   for any r within 2 FP64 ulp of 1 / d,

because the exact quotient of two binary32 numbers is either a binary32
number or at least ~2^-49 (relative) away from every binary32 rounding
midpoint, while fl64(n * r) is within ~2^-51 of it.  The device computes r
with v_rcp_f64 and two Newton steps; this checks the identity against
numpy's IEEE float32 division, with r perturbed by -2..+2 ulp, on random
operands over the walks' range of magnitudes, on quotients that are exact
binary32 numbers, on denormal results and on operands near the float
limits.  The device's results are pinned by the GPU parity tests (ORDERED
accelerations bit-exact with the oracle, which divides)."""

import numpy as np


def _fl64_quotient(n, d, ulps):
    r = 1.0 / d.astype(np.float64)
    for _ in range(abs(ulps)):
        r = np.nextafter(r, np.inf if ulps > 0 else -np.inf)
    with np.errstate(over="ignore", under="ignore"):
        return (n.astype(np.float64) * r).astype(np.float32)


def _check(n, d):
    with np.errstate(over="ignore", under="ignore"):
        q = n / d
    for ulps in (-2, -1, 0, 1, 2):
        got = _fl64_quotient(n, d, ulps)
        bad = got != q
        assert not np.any(bad), (f"{int(bad.sum())} mismatches at {ulps} ulp, e.g. n={n[bad][0]!r} "
                                 f"d={d[bad][0]!r}: {got[bad][0]!r} vs {q[bad][0]!r}")


def test_random_operands():
    rng = np.random.default_rng(0)
    for _ in range(4):
        k = 2_000_000
        sign = np.where(rng.random(k) < 0.5, -1.0, 1.0)
        n = (sign * rng.uniform(1, 2, k) * 2.0 ** rng.integers(-60, 40, k)).astype(np.float32)
        d = (rng.uniform(1, 2, k) * 2.0 ** rng.integers(-40, 30, k)).astype(np.float32)
        _check(n, d)


def test_exact_quotients():
    """n = q * d exactly in binary32 (q and d with few significant bits)."""
    rng = np.random.default_rng(1)
    k = 1_000_000
    q = (rng.integers(1, 1 << 11, k) * 2.0 ** rng.integers(-20, 20, k)).astype(np.float32)
    d = (rng.integers(1, 1 << 12, k) * 2.0 ** rng.integers(-20, 20, k)).astype(np.float32)
    n = (q.astype(np.float64) * d.astype(np.float64)).astype(np.float32)
    assert np.array_equal(n / d, q)
    _check(n, d)


def test_denormal_and_extreme_results():
    rng = np.random.default_rng(2)
    k = 1_000_000
    # quotients far below FLT_MIN (denormal or zero) and near FLT_MAX
    n = (rng.uniform(1, 2, k) * 2.0 ** rng.integers(-149, -120, k)).astype(np.float32)
    d = (rng.uniform(1, 2, k) * 2.0 ** rng.integers(0, 20, k)).astype(np.float32)
    _check(n, d)
    n = (rng.uniform(1, 2, k) * 2.0 ** rng.integers(100, 127, k)).astype(np.float32)
    d = (rng.uniform(1, 2, k) * 2.0 ** rng.integers(-30, 0, k)).astype(np.float32)
    _check(n, d)
    # denormal denominators
    n = (rng.uniform(1, 2, k) * 2.0 ** rng.integers(-140, -100, k)).astype(np.float32)
    d = (rng.uniform(1, 2, k) * 2.0 ** rng.integers(-149, -127, k)).astype(np.float32)
    _check(n, d)
