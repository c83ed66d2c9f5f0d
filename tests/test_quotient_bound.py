"""The identity the implicit FMM's criteria use instead of an FP32 division
(n-body_amd/csrc/fmm.hip, quotient_below): for FP32 x, y >= 0 and t >= 0,

    fl32(x / y) < t   <=>   x < m y   (exactly, in FP64),

with p the float below t and m = (p + t) / 2.  The device's results are pinned
by the GPU tests (equal interaction counts with the oracle, which divides);
this checks the identity itself against numpy's IEEE float32 division on
quotients spread tightly around t, on the candidates for ties, and on zeros
and infinities."""

import numpy as np

THETAS = (np.float32(0.4) * np.float32(0.4), np.float32(0.4), np.float32(0.5) ** 2, np.float32(0.25),
          np.float32(1.0), np.float32(0.7), np.float32(3.0e-3), np.float32(0.41289064) ** 2, np.float32(0.0))


def _mid(t):
    t = np.float32(t)
    p = np.nextafter(t, np.float32(-np.inf), dtype=np.float32)
    return (np.float64(p) + np.float64(t)) / 2.0


def _below(x, y, mid):
    return x.astype(np.float64) < mid * y.astype(np.float64)


def _ref(x, y, t):
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        return (x / y) < np.float32(t)


def test_quotient_bound_matches_fp32_division():
    rng = np.random.default_rng(5)
    for t in THETAS:
        mid = _mid(t)
        y = rng.uniform(1e-6, 1e3, 200000).astype(np.float32)
        x = (y.astype(np.float64) * np.float64(t) * rng.uniform(0.999, 1.001, y.size)).astype(np.float32)
        assert np.array_equal(_below(x, y, mid), _ref(x, y, t))
        # x one float either side of the nearest float to m y
        xm = (mid * y.astype(np.float64)).astype(np.float32)
        for xx in (xm, np.nextafter(xm, np.float32(np.inf)), np.nextafter(xm, np.float32(0))):
            assert np.array_equal(_below(xx, y, mid), _ref(xx, y, t))


def test_quotient_bound_edge_cases():
    x = np.array([0.0, 0.0, 1.0, np.inf, 1.0, 0.0, np.inf, 3.0e-45], np.float32)
    y = np.array([0.0, 1.0, 0.0, 1.0, np.inf, np.inf, np.inf, 1.0], np.float32)
    for t in THETAS:
        assert np.array_equal(_below(x, y, _mid(t)), _ref(x, y, t))
