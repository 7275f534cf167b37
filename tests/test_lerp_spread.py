"""The spread helper of the parity tests (tests/lerp.py) against a brute-force
scan: the largest minus the smallest log10 weight over curvature +- delta,
on a table with steep bins, flat bins and a non-monotone stretch."""
import numpy as np

from tests.lerp import joint_log10w, joint_log10w_spread


class _Table:
    def __init__(self):
        rng = np.random.default_rng(5)
        self.table = np.concatenate([np.linspace(1.0, 0.05, 60), np.full(20, 0.05), rng.uniform(0.01, 1.0, 20)]
                                    ).astype(np.float32)
        self.minK, self.maxK, self.step, self.ds = 0.0, 9.9, 0.1, 0.1


def test_spread_matches_a_scan():
    t = _Table()
    rng = np.random.default_rng(1)
    k = np.concatenate([rng.uniform(0.0, 9.9, 300), np.arange(1, 99) * 0.1, [0.0, 9.9, 10.5]])
    for delta in (1e-5, 1e-3, 0.25):
        lw, spread = joint_log10w_spread(k, t, 0.0, delta)
        assert np.array_equal(lw, joint_log10w(k, t, 0.0)[0])
        for kk, s in zip(k, spread):
            g = np.linspace(kk - delta, kk + delta, 4001)
            v = joint_log10w(g, t, 0.0)[0]
            scan = v.max() - v.min()
            # the scan's float32 points may round half an ulp past the ends
            assert s >= scan * (1 - 1e-2) - 1e-7, (kk, delta, s, scan)
            assert s <= scan + 1e-6 + 1e-2 * scan, (kk, delta, s, scan)
