"""The allocation-free passive-point entry point
(nbg_sys_accelerations_at_noalloc, static device scratch) against
nbg_sys_accelerations_at on the same system: bit-identical accelerations
and counts, on two calls in a row (the second with fewer targets, so stale
scratch contents from the first call would show), for both trees and both
sum modes, and a clear error above NBG_NOALLOC_MAX_POINTS.

Workload: Plummer sphere of 20,000 bodies (seed 3), passive targets
uniform in the cube [-1.5, 1.5]^3 (8,192 then 1,000 points, seeds 4 and 5).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NOALLOC_MAX_POINTS = 262144


def _targets(n, seed):
    r = np.random.default_rng(seed)
    return [r.uniform(-1.5, 1.5, n).astype(np.float32) for _ in range(3)]


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_equals_accelerations_at(nbg, gpu_ctx, solver, mode):
    b = nbg.plummer(20000, seed=3)
    p = nbg.default_params()
    p.solver, p.theta, p.leaf_size, p.morton_bits, p.eps2, p.sum_mode = solver, 0.5, 16, 63, 1e-4, mode
    s = nbg.System(gpu_ctx, p, b.n)
    s.set_state(b)
    print(f"[noalloc] solver {solver} mode {mode}", flush=True)
    for nt, seed in ((8192, 4), (1000, 5)):
        t = _targets(nt, seed)
        ref = s.accelerations_at(*t)
        ref_counts = s.counts()
        got = s.accelerations_at_noalloc(*t)
        assert s.counts() == ref_counts
        assert np.array_equal(got, ref), f"{nt} targets: allocation-free result differs"
        assert np.all(np.isfinite(got)) and np.any(got != 0.0)
    s.close()


def test_noalloc_refuses_too_many_targets(nbg, gpu_ctx):
    b = nbg.plummer(1000, seed=3)
    p = nbg.default_params()
    s = nbg.System(gpu_ctx, p, b.n)
    s.set_state(b)
    t = [np.zeros(NOALLOC_MAX_POINTS + 1, np.float32) for _ in range(3)]
    with pytest.raises(nbg.NbgError, match="NBG_NOALLOC_MAX_POINTS"):
        s.accelerations_at_noalloc(*t)
    s.close()
