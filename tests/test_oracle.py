"""Pins the CPU oracle (test infrastructure) against known answers, closed forms
and an independent NumPy formulation -- each check chosen so that a plausible
mistake (wrong range end, rotation sign, lerp normalisation, missing argument
swap, Simpson node set) fails it.
"""
import math

import numpy as np
import pytest

from tests import cases
from tests.lerp import joint_log10w

# ---------------------------------------------------------------------------
# A1 Philox4x32-10: published known-answer vectors (Random123 kat_vectors).
# ---------------------------------------------------------------------------
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(oracle, ctr, key, expect):
    assert list(oracle.philox(ctr, key)) == expect


@pytest.mark.parametrize("N", [4, 5, 6, 9, 27, 30])
def test_draw_ranges_exhaustive(oracle, N):
    """diff ~ U{2..min(N-2,25)}, left ~ U{1..N-diff-1}: every value, nothing else."""
    max_diff = min(N - 2, 25)
    seen = {}
    for path in range(3000 if N < 10 else 20000):
        d, l, x = oracle.draw(7, 0, path, path % 5, N)
        assert 2 <= d <= max_diff
        assert 1 <= l <= N - d - 1
        assert -0.5 < x < 0.5
        # x is an odd multiple of 2^-25: exactly representable, never 0
        q = x * 2.0 ** 25
        assert q == round(q) and int(round(q)) % 2 == 1
        seen.setdefault(d, set()).add(l)
    assert set(seen) == set(range(2, max_diff + 1))
    for d, lefts in seen.items():
        assert lefts == set(range(1, N - d))


def test_draw_uniformity(oracle):
    N, n = 12, 20000
    diffs = np.array([oracle.draw(3, 0, p, 0, N)[0] for p in range(n)])
    counts = np.bincount(diffs, minlength=11)[2:11]
    expect = n / 9
    chi2 = ((counts - expect) ** 2 / expect).sum()
    assert chi2 < 30.0  # 8 dof; p ~ 2e-4


def test_draw_streams_and_counters_differ(oracle):
    a = oracle.draw(1, 0, 5, 3, 64)
    assert oracle.draw(1, 0, 5, 3, 64) == a
    assert oracle.draw(1, 1, 5, 3, 64) != a
    assert oracle.draw(1, 0, 6, 3, 64) != a
    assert oracle.draw(2, 0, 5, 3, 64) != a


# ---------------------------------------------------------------------------
# A2/A3 quaternion rotation (CurvePerturbUtils.cpp:91-127).
# ---------------------------------------------------------------------------
def rodrigues(v, n, theta):
    v, n = np.asarray(v, float), np.asarray(n, float)
    return v * math.cos(theta) + np.cross(n, v) * math.sin(theta) + n * np.dot(n, v) * (1 - math.cos(theta))


def quat(n, theta):
    n = np.asarray(n, float)
    return np.array([math.cos(theta / 2), *(n * math.sin(theta / 2))])


def test_quaternion_quarter_turn(oracle):
    v = np.array([1.0, 0.0, 0.0])
    oracle.lib().oracle_rotate_vector_by_quaternion(quat([0, 0, 1], math.pi / 2), v)
    assert np.allclose(v, [0.0, 1.0, 0.0], atol=1e-15)  # active, counter-clockwise


def test_quaternion_matches_rodrigues(oracle):
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        th = rng.uniform(-math.pi, math.pi)
        v = rng.normal(size=3)
        w = v.copy()
        oracle.lib().oracle_rotate_vector_by_quaternion(quat(n, th), w)
        assert np.allclose(w, rodrigues(v, n, th), atol=1e-13)


def test_rotate_quat_by_quat_is_reversed_hamilton(oracle):
    a, b = np.array([0.3, -1.2, 0.7, 2.0]), np.array([-0.4, 0.5, 1.1, -0.2])

    def hamilton(p, q):
        return np.array([p[0] * q[0] - p[1] * q[1] - p[2] * q[2] - p[3] * q[3],
                         p[0] * q[1] + p[1] * q[0] + p[2] * q[3] - p[3] * q[2],
                         p[0] * q[2] - p[1] * q[3] + p[2] * q[0] + p[3] * q[1],
                         p[0] * q[3] + p[1] * q[2] - p[2] * q[1] + p[3] * q[0]])

    out = np.zeros(4)
    oracle.lib().oracle_rotate_quat_by_quat(a, b, out)
    assert np.allclose(out, hamilton(b, a), atol=1e-14)


def _arc(oracle, N, L):
    pos, alpha = oracle.initial_arc(N, L)
    return pos.copy(), alpha


def test_perturbation_invariants(oracle):
    N, L = 16, 1.3
    pos, _ = _arc(oracle, N, L)
    rng = np.random.default_rng(1)
    for _ in range(50):
        d = int(rng.integers(2, 15))
        left = int(rng.integers(1, N - d))
        x = np.float32(rng.uniform(-0.49, 0.49))
        before = pos.copy()
        flat = np.ascontiguousarray(pos.reshape(-1))
        oracle.lib().oracle_apply_perturbation_f64(flat, left, d, float(x))
        after = flat.reshape(N + 1, 3)
        right = left + d
        # points outside (left, right) do not move
        assert np.array_equal(after[: left + 1], before[: left + 1])
        assert np.array_equal(after[right:], before[right:])
        # rigid: every segment length preserved
        seg_b = np.linalg.norm(np.diff(before, axis=0), axis=1)
        seg_a = np.linalg.norm(np.diff(after, axis=0), axis=1)
        assert np.allclose(seg_a, seg_b, rtol=1e-12)
        # the interior moved as Rodrigues about p_left -> p_right by 2 pi x
        n = (before[right] - before[left]) / np.linalg.norm(before[right] - before[left])
        for k in range(left + 1, right):
            expect = before[left] + rodrigues(before[k] - before[left], n, 2 * math.pi * float(x))
            assert np.allclose(after[k], expect, atol=1e-12)
        # +theta then -theta restores the curve
        oracle.lib().oracle_apply_perturbation_f64(flat, left, d, float(-x))
        assert np.allclose(flat.reshape(N + 1, 3), before, atol=1e-12)
        pos = after


# ---------------------------------------------------------------------------
# A4/A5 tangents and joint curvature (CurvePerturbUtils.cpp:147-154, 219-243).
# ---------------------------------------------------------------------------
def _curv(oracle, pos, N, L):
    tan = np.zeros(N * 3)
    curv = np.zeros(N - 1)
    oracle.lib().oracle_tangents_curvatures_f64(np.ascontiguousarray(pos.reshape(-1)), N, L, tan, curv)
    return tan.reshape(N, 3), curv


def test_curvature_closed_forms(oracle):
    N, L = 40, 1.6
    inv_ds = float(np.float32(N) / np.float32(L))
    # straight line: 0
    line = np.zeros((N + 1, 3))
    line[:, 2] = np.linspace(0, L, N + 1)
    _, k = _curv(oracle, line, N, L)
    assert len(k) == N - 1 and np.allclose(k, 0.0, atol=1e-9)
    # full reversal at one joint: 2/ds
    rev = line.copy()
    rev[N // 2 + 1:, 2] = rev[N // 2, 2] - (rev[N // 2 + 1:, 2] - rev[N // 2, 2])
    _, k = _curv(oracle, rev, N, L)
    assert math.isclose(k[N // 2 - 1], 2.0 * inv_ds, rel_tol=1e-12)
    # regular arc: interior joints 2 sin(alpha/2) / ds
    pos, alpha = _arc(oracle, N, L)
    _, k = _curv(oracle, pos, N, L)
    assert np.allclose(k[1:N - 2], 2.0 * math.sin(alpha / 2) * inv_ds, rtol=1e-9)


# ---------------------------------------------------------------------------
# A11 initial arc (build-defined).
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N,L", [(20, 1.2), (64, 1.5), (100, 1.3), (1024, 2.0), (5, 1.8)])
def test_initial_arc(oracle, N, L):
    pos, alpha = _arc(oracle, N, L)
    ds = float(np.float32(L) / np.float32(N))
    seg = np.linalg.norm(np.diff(pos, axis=0), axis=1)
    assert np.allclose(seg, ds, rtol=1e-12)
    assert np.allclose(pos[0], [0, 0, 0]) and np.allclose(pos[N], [0, 0, 1])
    assert np.allclose(pos[1], [0, 0, ds]) and np.allclose(pos[N - 1], [0, 0, 1 - ds])
    # planar: every point in the plane spanned by the chord (z) and the bulge (x)
    assert np.allclose(pos[:, 1], 0.0, atol=1e-14)
    # closed form of the turning angle: |p_{N-1} - p_1| = ds sin(M a/2) / sin(a/2)
    M = N - 2
    assert math.isclose(ds * math.sin(M * alpha / 2) / math.sin(alpha / 2), 1 - 2 * ds, rel_tol=1e-10)


def test_arc_feasibility_matches_minimum_arclength(oracle):
    bc = oracle.bc12(edir=(0.6, 0.0, 0.8))
    N = 30
    lmin = oracle.lib().oracle_minimum_arclength(bc, N)
    pos = np.zeros((N + 1) * 3)
    a = np.zeros(1)
    import ctypes as C
    alpha = C.c_double()
    assert oracle.lib().oracle_initial_arc(bc, float(lmin) * 1.001, N, pos, C.byref(alpha)) == 0
    assert oracle.lib().oracle_initial_arc(bc, float(lmin) * 0.999, N, pos, C.byref(alpha)) == 1
    del a


# ---------------------------------------------------------------------------
# A6 per-path weight (PathWeighters.cpp:27-104).
# ---------------------------------------------------------------------------
class _T:
    def __init__(self, table, ds, minK, maxK, step):
        self.table = np.asarray(table, np.float32)
        self.ds, self.minK, self.maxK, self.step = ds, minK, maxK, step


def test_weight_constant_table(oracle):
    c, a, ds = 0.37, 0.25, 0.05
    t = _T(np.full(64, c), ds, 0.0, 31.5, 0.5)
    k = np.linspace(0.0, 31.0, 17).astype(np.float32)
    ok, w = oracle.weight_curvatures(k, t, a)
    absorb = float(np.float32(math.exp(float(np.float32(-np.float32(a) * np.float32(ds))))))
    # float lerp of equal entries rounds at 1e-7; a missing absorption term is off by 0.09
    assert ok and math.isclose(w, len(k) * math.log10(np.float32(c) * absorb), rel_tol=1e-6)


def test_weight_lerp_in_curvature_units(oracle):
    """T[i] = i + 1, step 0.5: kappa 1.2 -> bin 2, fraction 0.2 (not 0.4)."""
    t = _T(np.arange(1, 65), 0.1, 0.0, 31.5, 0.5)
    ok, w = oracle.weight_curvatures(np.array([1.2], np.float32), t, 0.0)
    assert ok and math.isclose(10 ** w, 3.0 * 0.8 + 4.0 * 0.2, rel_tol=1e-6)
    # last bin uses itself on the right (PathWeighters.cpp:72-74) and clamps above maxK
    ok, w = oracle.weight_curvatures(np.array([31.5, 40.0], np.float32), t, 0.0)
    assert ok and math.isclose(10 ** w, 64.0 * 64.0, rel_tol=1e-6)


def test_weight_nan_and_zero(oracle):
    t = _T(np.ones(16), 0.1, 0.0, 7.5, 0.5)
    t.table[3] = np.nan
    ok, _ = oracle.weight_curvatures(np.array([1.6], np.float32), t, 0.0)
    assert not ok
    t.table[:] = 0.0
    ok, w = oracle.weight_curvatures(np.array([1.0], np.float32), t, 0.0)
    assert ok and w == -math.inf


# ---------------------------------------------------------------------------
# A7 table (PathWeightUtils.cpp:19-33, 281-433).
# ---------------------------------------------------------------------------
def test_gaussian_phase_closed_form(oracle):
    for m in (0.0025, 0.01, 0.075):
        for p in (0.0, 1.0, 7.5):
            expect = math.sqrt(math.pi * m / 2) / (1 - math.exp(-2 / m)) * math.exp(-m * p * p / 2)
            assert math.isclose(oracle.lib().oracle_gaussian_phase(p, m), expect, rel_tol=2e-6)


@pytest.mark.parametrize("N,L,c", [(64, 1.5, 1.0), (20, 1.2, 1.0), (30, 3.0, 0.7)])
def test_table_scatter_free_closed_form(oracle, N, L, c):
    """scatter = 0: T(k) = bias/(2 pi^2) sqrt(pi/2) c^-3 exp(-(k ds)^2/(2 c^2)),
    with c the config `mu` (the argument swap makes mu the regulariser)."""
    ds = float(np.float32(L) / np.float32(N))
    t = oracle.build_table(ds, mu=c, eps=0.0025, scatter=0.0, absorption=0.0, numStepsInt=20000,
                           numCurvatureSteps=512)
    k = t.minK + np.arange(512, dtype=np.float32) * np.float32(t.step)
    kds = (k.astype(np.float32) * np.float32(ds)).astype(np.float64)
    expect = 1 / (2 * math.pi ** 2) * math.sqrt(math.pi / 2) * c ** -3 * np.exp(-kds ** 2 / (2 * c * c))
    rel = np.abs(t.table / expect - 1)
    assert rel.max() < 2e-6


def test_table_simpson_node_set(oracle):
    """numStepsInt = 4 keeps nodes 0 (w), 1 (4w) and 4 (w) only: with a p^2
    integrand (mu = 0, scatter = 0, kappa = 0) T(0) = (20 h^3 / 3) / (2 pi^2),
    not the full Simpson value 64 h^3 / 3."""
    hmax = np.float32(2.0)
    h = float(np.float32(hmax / np.float32(4)))
    v = oracle.lib().oracle_integrate(0.0, 0.0, 1.0, 0.0, 1.0, 4, 0.0, float(hmax), 0.1)
    expect = (20 * h ** 3 / 3) / (2 * float(np.float32(math.pi)) ** 2)
    assert math.isclose(v, expect, rel_tol=1e-6)


def test_table_range(oracle):
    for N, L in [(20, 1.2), (64, 1.5), (1024, 2.0)]:
        ds = np.float32(L) / np.float32(N)
        mn, mx, st = oracle.table_range(4096, float(ds))
        assert mn == 0.0
        assert mx == np.float32((2.0 / float(ds)) * float(np.float32(1.01)))
        assert st == np.float32(np.float32(mx) / np.float32(4095))


# ---------------------------------------------------------------------------
# A8 accumulator (CombinedWeightUtils.h:48-174).
# ---------------------------------------------------------------------------
def _lse10(w):
    w = np.asarray(w, np.longdouble)
    m = w.max()
    return float(m + np.log10(np.sum(np.power(np.longdouble(10), w - m))))


def test_accumulator_brute_force(oracle):
    rng = np.random.default_rng(4)
    w = rng.uniform(-120.0, -60.0, size=2_500_000)  # crosses the 1e6-path accumulators
    got = oracle.lib().oracle_accumulate_log10(w, len(w))
    assert math.isclose(got, _lse10(w), rel_tol=0, abs_tol=2e-12)
    perm = rng.permutation(len(w))
    assert math.isclose(oracle.lib().oracle_accumulate_log10(np.ascontiguousarray(w[perm]), len(w)), got,
                        abs_tol=2e-12)


def test_accumulator_combine_equals_add(oracle):
    rng = np.random.default_rng(5)
    w = rng.uniform(-50.0, 10.0, size=4000)
    one = oracle.lib().oracle_accumulate_log10(w, len(w))
    two = oracle.lib().oracle_accumulate_combine_log10(w, 1234, len(w))
    assert math.isclose(one, two, abs_tol=1e-12)


def test_accumulator_first_zero_weight_is_nan(oracle):
    """The reference's AddValue with -inf as the first value sets offset = +inf
    and the running total to NaN (CombinedWeightUtils.h:62-69); the oracle keeps
    that, and the exact log-sum-exp aggregate is reported beside it."""
    w = np.array([-math.inf, -3.0, -4.0])
    assert math.isnan(oracle.lib().oracle_accumulate_log10(w, 3))
    w = np.array([-3.0, -math.inf, -4.0])
    assert math.isclose(oracle.lib().oracle_accumulate_log10(w, 3), _lse10([-3.0, -4.0]), abs_tol=1e-12)


# ---------------------------------------------------------------------------
# A9 path normaliser (PathWeightUtils.cpp:643-752).
# ---------------------------------------------------------------------------
def test_norm_k4_and_default(oracle):
    bc = oracle.bc12()
    ds = 0.4
    z = abs(1.0 / ds - 2.0)
    expect = (1.0 / z) * 2 * math.pi / ds ** 3 if z < 2 else 0.0
    assert math.isclose(oracle.lib().oracle_norm(4, ds, bc, 1.6, 2, 2), expect, rel_tol=1e-5)
    for N in (7, 20, 64):
        assert oracle.lib().oracle_norm(N, 0.05, bc, 1.0, 2, 2) == 1.0


def random_flight_density(n, r):
    """3-D density of the sum of n independent uniform unit vectors at distance
    r (Rayleigh 1919; Treloar 1946): a piecewise polynomial in r over r,
    1 / (2^(n+1) pi (n-2)! r) sum_k (-1)^k C(n,k) (n - r - 2k)^(n-2), 0 < r < n."""
    if r >= n:
        return 0.0
    s = sum((-1) ** k * math.comb(n, k) * (n - r - 2 * k) ** (n - 2) for k in range(int((n - r) // 2) + 1))
    return s / (2 ** (n + 1) * math.pi * math.factorial(n - 2) * r)


def test_random_flight_density_closed_form():
    """The closed form used below, pinned by what it must satisfy: unit mass for
    every n, the two-step density 1/(8 pi r) on (0, 2) (|s1 + s2|^2 = 2 + 2 cos t
    with cos t uniform), and p_{n+1} = p_n convolved with the unit sphere's
    uniform measure (checked numerically at n = 3 -> 4)."""
    for n in (2, 3, 4, 5, 6):
        r = np.linspace(1e-6, n, 400001)
        mass = np.trapezoid([4 * math.pi * x * x * random_flight_density(n, x) for x in r[::50]], r[::50])
        assert abs(mass - 1.0) < 2e-4, (n, mass)
    for r in (0.1, 0.7, 1.3, 1.99):
        assert math.isclose(random_flight_density(2, r), 1.0 / (8 * math.pi * r), rel_tol=1e-12)
    # sphere convolution: (f * sigma)(w) / (4 pi) = 1/2 int_{-1}^{1} f(sqrt(w^2 + 1 - 2 w u)) du
    u = np.linspace(-1.0, 1.0, 200001)
    for w in (0.3, 1.2, 2.5, 3.4):
        rr = np.sqrt(w * w + 1.0 - 2.0 * w * u)
        conv = 0.5 * np.trapezoid([random_flight_density(3, x) for x in rr], u)
        assert math.isclose(conv, random_flight_density(4, w), rel_tol=2e-4), (w, conv)


def _bc_for_z(D):
    """Boundary (0,0,0) -> (0,0,D), both tangents +z: the normaliser's
    argument is z = |D/ds - 2| (PathWeightUtils.cpp:730-734), computed in float."""
    return np.array([0, 0, 0, 0, 0, 1, 0, 0, D, 0, 0, 1], dtype=np.float64)


def _z_of(D, ds):
    f = np.float32
    return float(abs(f(f(D) * (f(1.0) / f(ds))) - f(1.0) - f(1.0)))


def test_norm_nested_quadrature_converges_to_random_flight(oracle):
    """N = 5, 6 (PathWeightUtils.cpp:653-752).  K4 = 2 pi / (z ds^3) on z < 2 is
    (4 pi)^2 p_2(z) / ds^3, and each F(N, M) / K(N) level is a left Riemann rule
    for a convolution with the unit sphere (2 pi int f(|w - s|) d cos).  The
    K(N) macro nests three levels over K4 for every N, so K5 = K6 ->
    (4 pi)^5 p_5(z) / ds^3.  The oracle converges to that at first order in the
    node spacing (the error halves as the node counts double) at every z, and
    the Richardson combination 2 v(2n) - v(n) lands within 5e-3.  A missing or
    extra level, a lost 2 pi or a wrong node spacing converges elsewhere."""
    ds = 0.4
    L = oracle.lib()
    for D in (1.0, 1.4, 1.8, 2.2, 2.6):
        z = _z_of(D, ds)
        closed = (4 * math.pi) ** 5 * random_flight_density(5, z) / ds ** 3
        v = {n: L.oracle_norm(5, ds, _bc_for_z(D), 5.0, n, n) for n in (100, 200, 400)}
        e = {n: (v[n] - closed) / closed for n in v}
        assert abs(e[400]) < 0.07, (z, e)
        assert 1.7 < e[200] / e[400] < 2.3 and 1.7 < e[100] / e[200] < 2.3, (z, e)
        assert abs((2 * v[400] - v[200]) - closed) / closed < 5e-3, (z, e)
        assert L.oracle_norm(6, ds, _bc_for_z(D), 5.0, 100, 100) == v[100]  # K6 is K5
    # beyond the reach of five unit steps the density, and the quadrature, vanish
    assert L.oracle_norm(5, ds, _bc_for_z(3.0), 5.0, 60, 60) == 0.0


# ---------------------------------------------------------------------------
# End to end: an independent NumPy path (Rodrigues on positions, float32 lerp)
# against the oracle's FP64 path for small N.
# ---------------------------------------------------------------------------
def numpy_paths(oracle, pos0, N, L, K, n, table, absorption, seed=1, chain_length=1, burn_in=0):
    """Chains of chain_length paths; each chain first applies burn_in unweighted
    paths (the skipped paths of FullExperimentRunnerOptimalPerturb.cpp:559-561)
    drawn from path ids 2^63 | (chain * burn_in + b)."""
    out = []
    inv_ds = np.float32(N) / np.float32(L)

    def perturb(pos, path):
        for k in range(K):
            d, l, x = oracle.draw(seed, 0, path, k, N)
            r = l + d
            axis = (pos[r] - pos[l]) / np.linalg.norm(pos[r] - pos[l])
            for i in range(l + 1, r):
                pos[i] = pos[l] + rodrigues(pos[i] - pos[l], axis, 2 * math.pi * float(x))

    for p in range(n):
        if p % chain_length == 0:
            pos = pos0.copy()
            for b in range(burn_in):
                perturb(pos, (1 << 63) | ((p // chain_length) * burn_in + b))
        perturb(pos, p)
        t = np.diff(pos, axis=0)
        t /= np.linalg.norm(t, axis=1, keepdims=True)
        curv = np.linalg.norm(np.diff(t, axis=0) * float(inv_ds), axis=1)
        lw, _ = joint_log10w(curv[None, :], table, absorption)
        out.append(lw.sum())
    return np.array(out)


@pytest.mark.parametrize("case", [cases.SMALL, cases.MID], ids=lambda c: c.name)
def test_end_to_end_against_numpy(oracle, case):
    table, pos32 = cases.setup(case, oracle)
    n = 64
    ref = oracle.simulate(pos32.astype(np.float64), case.N, case.L, case.K, n, table, case.absorption,
                          seed=case.seed)
    npw = numpy_paths(oracle, pos32.astype(np.float64), case.N, case.L, case.K, n, table, case.absorption,
                      seed=case.seed)
    assert np.allclose(ref["log10w"], npw, rtol=1e-9, atol=1e-9)


def test_chain_mode_replays_from_chain_start(oracle):
    case = cases.MID
    table, pos32 = cases.setup(case, oracle)
    full = oracle.simulate(pos32.astype(np.float64), case.N, case.L, 2, 96, table, case.absorption, chain_length=32)
    part = oracle.simulate(pos32.astype(np.float64), case.N, case.L, 2, 40, table, case.absorption, chain_length=32,
                           path_offset=40)
    assert np.array_equal(full["log10w"][40:80], part["log10w"])


def test_burn_in_against_numpy(oracle):
    """Burn-in (pathsToSkip) against the NumPy chain restatement; burn-in 0 is
    the plain chain; a range split replays the burn-in of its chain."""
    case = cases.MID
    table, pos32 = cases.setup(case, oracle)
    pos = pos32.astype(np.float64)
    ref = oracle.simulate(pos, case.N, case.L, 2, 48, table, case.absorption, seed=3, chain_length=16, burn_in=5)
    npw = numpy_paths(oracle, pos, case.N, case.L, 2, 48, table, case.absorption, seed=3, chain_length=16, burn_in=5)
    assert np.allclose(ref["log10w"], npw, rtol=1e-9, atol=1e-9)
    plain = oracle.simulate(pos, case.N, case.L, 2, 48, table, case.absorption, seed=3, chain_length=16)
    assert np.array_equal(plain["log10w"], oracle.simulate(pos, case.N, case.L, 2, 48, table, case.absorption,
                                                           seed=3, chain_length=16, burn_in=0)["log10w"])
    assert not np.allclose(plain["log10w"][::16], ref["log10w"][::16])
    part = oracle.simulate(pos, case.N, case.L, 2, 20, table, case.absorption, seed=3, chain_length=16, burn_in=5,
                           path_offset=20)
    assert np.array_equal(ref["log10w"][20:40], part["log10w"])


def test_reference_fp32_close_to_fp64(oracle):
    """The reference's own float positions stay within 1e-5 per path at C2."""
    case = cases.C2
    a = cases.oracle_run(case, oracle, 512)
    b = cases.oracle_run(case, oracle, 512, precision=32)
    rel = np.abs(a["log10w"] - b["log10w"]) / np.abs(a["log10w"])
    assert rel.max() < 2e-5


# ---------------------------------------------------------------------------
# Independent-path estimator (oracle_generate.cpp): ExperimentBase.cpp:617-983,
# 1246-1386.
# ---------------------------------------------------------------------------
def _gen_levels_py(N):
    n0 = N - 2
    for k in range(1, 20):
        if n0 in (2 << k, 3 << k):
            return k, n0 >> k
    return None


def test_generate_segment_count_rule(oracle):
    """N - 2 must halve down to 2 or 3 (ExperimentBase.cpp:810-817), exhaustively."""
    for N in range(1, 600):
        assert oracle.generate_levels(N) == _gen_levels_py(N), N


def _words(oracle, seed, path, node, sub=0):
    r = oracle.philox([node, 0x80000000 | sub, path & 0xFFFFFFFF, path >> 32], [seed & 0xFFFFFFFF, seed >> 32])
    return [(int(r[0]) >> 8) / 2.0 ** 24, (int(r[1]) >> 8) / 2.0 ** 24, int(r[2]) & 1, (int(r[3]) >> 8) / 2.0 ** 24]


def _gen_setup(oracle, N, L, bc):
    ds = float(np.float32(L) / np.float32(N))
    t = oracle.build_table(ds, eps=0.0025, numStepsInt=2000, numCurvatureSteps=1024)
    return ds, t


GEN_CASES = [(6, 1.4, (0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1)), (8, 1.7, (0, 0, 0), (1, 0, 0), (0.5, 0.3, 0.2),
                                                                    (0, 1, 0)),
             (14, 1.4, (0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1)), (34, 2.5, (0, 0, 0), (0, 0, 1), (1, 0, 0.2),
                                                                     (0.6, 0.8, 0.0)),
             (66, 1.6, (0, 0, 0), (0, 0, 1), (0, 0, 1), (0, 0, 1))]


@pytest.mark.parametrize("N,L,s,sd,e,ed", GEN_CASES)
def test_generated_paths_structure(oracle, N, L, s, sd, e, ed):
    """Every generated path: fixed start, end and hinge points; each leaf's
    segments exactly ds unless the leaf is within the reference's 0.001
    straight tolerance (then its point is the midpoint); each centre within
    reach (perSide * ds) of both ends of its span."""
    bc = oracle.bc12(s, np.array(sd) / np.linalg.norm(sd), e, np.array(ed) / np.linalg.norm(ed))
    ds, t = _gen_setup(oracle, N, L, bc)
    r = oracle.generate(bc, L, N, 300, t, 0.004, seed=5, want_pos=True)
    assert r["valid"].all()
    P = r["positions"]
    k, leaf = oracle.generate_levels(N)
    n0 = N - 2
    assert np.array_equal(P[:, 0], np.broadcast_to(bc[0:3], (300, 3)))
    assert np.array_equal(P[:, N], np.broadcast_to(bc[6:9], (300, 3)))
    assert np.abs(P[:, 1] - (bc[0:3] + ds * bc[3:6])).max() < 1e-15
    assert np.abs(P[:, N - 1] - (bc[6:9] - ds * bc[9:12])).max() < 1e-15
    exact_leaves = 0
    for lv in range(k):
        span = n0 >> lv
        for q in range(1 << lv):
            lft = 1 + q * span
            c, a, b = P[:, lft + span // 2], P[:, lft], P[:, lft + span]
            reach = ds * (span // 2) + 1e-12
            assert (np.linalg.norm(c - a, axis=1) <= reach).all() and (np.linalg.norm(b - c, axis=1) <= reach).all()
    for q in range(1 << k):
        lft = 1 + q * leaf
        a, b = P[:, lft + leaf - 2], P[:, lft + leaf]
        m = P[:, lft + leaf - 1]
        d = np.linalg.norm(b - a, axis=1)
        straight = np.abs(2 * ds - d) < 0.001
        exact_leaves += int((~straight).sum())
        s1, s2 = np.linalg.norm(m - a, axis=1), np.linalg.norm(b - m, axis=1)
        assert np.abs(s1[~straight] - ds).max(initial=0) < 1e-12 and np.abs(s2[~straight] - ds).max(initial=0) < 1e-12
        assert np.abs(m[straight] - 0.5 * (a[straight] + b[straight])).max(initial=0) < 1e-15
    assert exact_leaves > 0


@pytest.mark.parametrize("N,L,s,sd,e,ed", GEN_CASES[:4])
def test_generated_joint_outputs(oracle, N, L, s, sd, e, ed):
    """The curvatures and branch margins the GPU parity rule reads: curv is the
    float of |t_{i+1} - t_i| N / L over the returned points' unit tangents, and
    margin is at most the distance of each leaf's and centre's 0.001 tolerance
    test (|span ds - d| against 0.001, d against 0.001), recomputed from the
    returned points, from its threshold; a path whose margin sits on one of
    these distances equals it."""
    bc = oracle.bc12(s, np.array(sd) / np.linalg.norm(sd), e, np.array(ed) / np.linalg.norm(ed))
    ds, t = _gen_setup(oracle, N, L, bc)
    n = 300
    r = oracle.generate(bc, L, N, n, t, 0.004, seed=5, want_pos=True, want_joints=True)
    P = r["positions"]
    seg = np.diff(P, axis=1)
    tan = seg / np.linalg.norm(seg, axis=2, keepdims=True)
    kap = np.linalg.norm(np.diff(tan, axis=1), axis=2) * float(np.float32(N) / np.float32(L))
    v = r["valid"]
    assert v.all()
    assert np.abs(r["curv"][v].astype(np.float64) - kap[v]).max() <= 1e-6 * max(1.0, kap.max())
    k, leaf = oracle.generate_levels(N)
    n0 = N - 2
    dist = []
    for lv in range(k):
        span = n0 >> lv
        for q in range(1 << lv):
            lft = 1 + q * span
            d = np.linalg.norm(P[:, lft + span] - P[:, lft], axis=1)
            dist += [np.abs(np.abs(span * ds - d) - 0.001), np.abs(d - 0.001)]
    for q in range(1 << k):
        lft = 1 + q * leaf
        d = np.linalg.norm(P[:, lft + leaf] - P[:, lft], axis=1)
        dist += [np.abs(np.abs(leaf * ds - d) - 0.001), np.abs(d - 0.001)]
    lowest = np.min(np.stack(dist), axis=0)
    m = r["margin"]
    assert (m > 0).all() and (m <= lowest + 1e-12).all()
    assert np.isclose(m, lowest, rtol=0, atol=1e-12).any()


def test_generated_two_segment_angle_is_the_drawn_theta(oracle):
    """N = 6: the leaf point sits on the circle of radius sqrt(ds^2 - d^2/4)
    around its chord at the angle 2 pi u1 of its node's Philox words, measured
    from m = normalize(n x e) towards n x m (ExperimentBase.cpp:672-697); the
    centre's polar angle from its span start satisfies cos(phi) = 1 - 2 u0 vmax,
    vmax = (1 - d / (2 rps)) / 2 (:910-935)."""
    N, L = 6, 1.5
    bc = oracle.bc12()
    ds, t = _gen_setup(oracle, N, L, bc)
    r = oracle.generate(bc, L, N, 64, t, 0.004, seed=9, want_pos=True)
    P = r["positions"]
    for p in range(64):
        # centre (node 1), span 4 between points 1 and 5
        a, b, c = P[p, 1], P[p, 5], P[p, 3]
        u0, u1, coin, u3 = _words(oracle, 9, p, 1)
        d = np.linalg.norm(b - a)
        z = (b - a) / d
        vmax = 0.5 * (1 - d / (2 * 2 * ds))
        rel = (c - a) if coin == 0 else (c - b)
        cosphi = np.dot(rel, z) / np.linalg.norm(rel) * (1 if coin == 0 else -1)
        assert cosphi == pytest.approx(1 - 2 * u0 * vmax, abs=1e-12)
        # leaves (nodes 2, 3)
        for q, (lft, rgt) in enumerate(((1, 3), (3, 5))):
            a2, b2, m2 = P[p, lft], P[p, rgt], P[p, lft + 1]
            d2 = np.linalg.norm(b2 - a2)
            if abs(2 * ds - d2) < 0.001:
                continue
            n = (b2 - a2) / d2
            e = np.array([0.0, 1.0, 0.0]) if abs(n[0]) >= 0.99 else np.array([1.0, 0.0, 0.0])
            m = np.cross(n, e)
            m /= np.linalg.norm(m)
            v = m2 - 0.5 * (a2 + b2)
            ang = math.atan2(np.dot(v, np.cross(n, m)), np.dot(v, m)) % (2 * math.pi)
            theta = 2 * math.pi * _words(oracle, 9, p, 2 + q)[1]
            assert min(abs(ang - theta), 2 * math.pi - abs(ang - theta)) < 1e-9


def test_generated_straight_line_closed_form(oracle):
    """Arclength equal to the end distance: every node takes the straight
    branch, the path is the segment from start to end, every curvature is 0 and
    the weight is (N - 1) log10(T[0] exp(-absorption ds))."""
    N = 10
    bc = oracle.bc12()
    ds, t = _gen_setup(oracle, N, 1.0, bc)
    r = oracle.generate(bc, 1.0, N, 16, t, 0.004, seed=2, want_pos=True)
    P = r["positions"]
    assert np.abs(P[:, :, :2]).max() == 0.0
    assert np.abs(P[:, :, 2] - np.linspace(0, 1, N + 1)).max() < 1e-7
    want = (N - 1) * math.log10(float(t.table[0]) * float(np.float32(math.exp(float(np.float32(-0.004 * np.float32(ds)))))))
    assert np.allclose(r["log10w"], want, rtol=1e-7)


def test_generated_multi_ds_table_choice(oracle):
    """Multi-ds form (ExperimentBase.cpp:1388-1597): with flat tables of
    distinct constants the path weight names the table, which must be the one
    whose ds is nearest to u (maxDs - minDs) + minDs in float (u from node 0's
    words), first on ties (PathWeightUtils.cpp:223-236)."""
    N, T, bins = 10, 7, 64
    dss = np.linspace(0.09, 0.15, T).astype(np.float32)
    consts = np.float32(1.0) + np.arange(T, dtype=np.float32)
    tables = np.repeat(consts[:, None], bins, axis=1)
    meta = np.stack([dss, np.zeros(T, np.float32), np.full(T, 100.0, np.float32), np.full(T, 100.0 / (bins - 1),
                                                                                        np.float32)], axis=1)
    lo, hi = np.float32(0.1), np.float32(0.14)
    r = oracle.generate_multi(oracle.bc12(), N, 200, tables, meta, 0.004, float(lo), float(hi), seed=4)
    assert r["valid"].all()
    for p in range(200):
        u = np.float32(_words(oracle, 4, p, 0)[0])
        # fmaf(u, hi - lo, lo): the product of two floats is exact in double
        ds = np.float32(float(u) * float(hi - lo) + float(lo))
        d = np.abs(dss - ds)
        t = int(np.argmin(d))
        absorb = float(np.float32(math.exp(float(np.float32(np.float32(-0.004) * dss[t])))))
        want = (N - 1) * math.log10(float(consts[t]) * absorb)
        assert r["log10w"][p] == pytest.approx(want, rel=1e-6), p  # float lerp and product per joint


def test_chain_log_is_the_natural_log(oracle):
    """The build-defined log of the chains' ds draw (frexp, atanh series) is
    the natural log to a few ulps over (0, 1]."""
    ys = np.concatenate([np.geomspace(1e-300, 1.0, 2000), 1.0 - np.geomspace(1e-16, 0.5, 500), [1.0, 0.5, 0.75]])
    for y in ys:
        want = math.log(y)
        got = oracle.bd_log(float(y))
        assert abs(got - want) <= 4e-16 * abs(want) + 1e-300, (y, got, want)


@pytest.mark.parametrize("t,max_ds", [(0.104, 0.03), (5.0, 1.0), (40.0, 0.2)])
def test_chain_ds_is_the_truncated_exponential(oracle, t, max_ds):
    """ds = minDs + x with x exponential of rate t truncated to [0, maxDs]
    (ExperimentBase.cpp:1662-1667 draws the same distribution by rejection):
    the draw inverts the truncated distribution function
    F(x) = (1 - exp(-t x)) / (1 - exp(-t maxDs)), so F(ds(u) - minDs) = u up to
    the float rounding of ds, ds(0) = minDs and ds rises with u."""
    t32, lo, hi = np.float32(t), np.float32(0.02), np.float32(max_ds)
    us = np.concatenate([np.arange(0, 1 << 24, 4099) / float(1 << 24), [1.0 - 2.0 ** -24]])
    ds = np.array([oracle.chain_ds(u, float(t32), float(lo), float(hi)) for u in us], np.float64)
    assert ds[0] == float(lo) and np.all(np.diff(ds) >= 0.0) and ds[-1] <= float(lo) + float(hi) * (1 + 1e-6)
    x = ds - float(lo)
    F = -np.expm1(-float(t32) * x) / -math.expm1(-float(t32) * float(hi))
    # one float ulp of ds moves F by at most the density times that ulp
    dens = float(t32) * np.exp(-float(t32) * x) / -math.expm1(-float(t32) * float(hi))
    slack = dens * np.spacing(ds.astype(np.float32)).astype(np.float64) + 1e-12
    assert np.all(np.abs(F - us) <= slack), np.max(np.abs(F - us) - slack)


def test_chain_sampling_weight_closed_form(oracle):
    """log10 of 1 / (t exp(-t (ds - minDs) N)), formed in float
    (ExperimentBase.cpp:1669-1670), against -log10 t + t (ds - minDs) N log10 e."""
    for t, ds, lo, N in [(0.104, 0.05, 0.03, 34), (1.5, 0.2, 0.1, 66), (0.01, 0.031, 0.03, 10)]:
        got = oracle.chain_log10_sampling_weight(t, ds, lo, N)
        t32, d32 = float(np.float32(t)), float(np.float32(np.float32(ds) - np.float32(lo)))
        want = -math.log10(t32) + t32 * d32 * N / math.log(10.0)
        assert got == pytest.approx(want, rel=1e-6, abs=1e-6)


def _chain_case(oracle, cache_root, n_chains=48, length=40, norm=2.0, seed=3):
    from twisty_amd import host
    N = 34
    bc13 = host.boundary13((0, 0, 0), (0, 0, 1), (0.4, 0, 1), (0.6, 0, 0.8), 2.0)
    bc13[9:12] /= np.linalg.norm(bc13[9:12])
    lo = np.float32(host.minimum_arclength(bc13, N) / N)
    hi = np.float32(0.03)
    tabs, _, _ = host.cached_rt_tables(float(lo) - 0.001, float(lo + hi) + 0.001, 8, 1.0,
                                       float(np.float32(0.08) ** 2), 0.1, 0.004, numStepsInt=2000,
                                       numCurvatureSteps=512, cache_root=cache_root)
    tables = np.stack([t.table for t in tabs])
    meta = np.array([[t.ds, t.minK, t.maxK, t.step] for t in tabs], np.float32)
    ext = float(np.float32(np.float32(0.004) + np.float32(0.1)))
    r = oracle.generate_chains(bc13[:12].astype(np.float64), N, n_chains, length, tables, meta, ext, float(lo),
                               float(hi), norm_log10=norm, seed=seed)
    return r, n_chains, length, norm


def test_chain_acceptance_rule(oracle, tmp_path):
    """The chains against a replay of ExperimentBase.cpp:1730-1784 from the
    oracle's own proposal weights and draws: an invalid proposal is skipped, a
    chain's first valid proposal is taken, a later one when
    u <= 10^(w + norm) / 10^(w' + norm), and every valid proposal adds the
    chain's current weight; the result is the sum over those.  Also: a
    proposal at least as heavy as the current state is always taken, and the
    current state is always one of the chain's own proposals."""
    r, C, L, norm = _chain_case(oracle, tmp_path)
    w, v, u, ch = (r[k].reshape(C, L) for k in ("log10w", "valid", "u", "chain"))
    total, nv = -math.inf, 0
    tests = taken = 0
    for c in range(C):
        cur = None
        for s in range(L):
            if v[c, s]:
                if cur is None:
                    cur = w[c, s] + norm
                else:
                    tests += 1
                    ratio = 10.0 ** (w[c, s] + norm) / 10.0 ** cur
                    if float(np.float32(u[c, s])) <= ratio:
                        cur = w[c, s] + norm
                        taken += 1
                    if w[c, s] + norm >= cur:
                        assert ch[c, s] + norm == w[c, s] + norm
                nv += 1
                m = max(total, cur)
                total = m + math.log10(10.0 ** (total - m) + 10.0 ** (cur - m))
            if cur is None:
                assert np.isnan(ch[c, s])
            else:
                assert ch[c, s] == pytest.approx(cur - norm, abs=1e-12), (c, s)
                assert np.any(np.abs(w[c, :s + 1] - ch[c, s]) <= 1e-12)
    assert nv == r["n_valid"] and 0 < taken < tests
    assert r["log10_sum_w"] == pytest.approx(total, rel=1e-12)


def test_chain_proposals_with_flat_tables(oracle):
    """Each proposal of the chains (ExperimentBase.cpp:1658-1718): with flat
    tables of distinct constants its weight names its table, which must be the
    one nearest the proposal's ds drawn from node 0's first word, weighted with
    absorption + scatter on every joint (PathWeighters.cpp:105-177) plus log10
    of the sampling weight; u is node 0's second word."""
    N, T, bins, C, L = 10, 7, 64, 20, 10
    dss = np.linspace(0.09, 0.16, T).astype(np.float32)
    consts = np.float32(1.0) + np.arange(T, dtype=np.float32)
    tables = np.repeat(consts[:, None], bins, axis=1)
    meta = np.stack([dss, np.zeros(T, np.float32), np.full(T, 100.0, np.float32), np.full(T, 100.0 / (bins - 1),
                                                                                        np.float32)], axis=1)
    lo, hi = np.float32(0.1), np.float32(0.05)
    ext = float(np.float32(np.float32(0.004) + np.float32(0.1)))
    r = oracle.generate_chains(oracle.bc12(), N, C, L, tables, meta, ext, float(lo), float(hi), seed=6, chain_offset=3)
    assert r["valid"].all()
    for c in range(C):
        for s in range(L):
            p, path = c * L + s, (3 + c) * L + s
            wd = _words(oracle, 6, path, 0)
            assert r["u"][p] == wd[1]
            ds = np.float32(oracle.chain_ds(wd[0], ext, float(lo), float(hi)))
            t = int(np.argmin(np.abs(dss - ds)))
            assert r["table"][p] == t
            absorb = float(np.float32(math.exp(float(np.float32(np.float32(-ext) * dss[t])))))
            want = (N - 1) * math.log10(float(consts[t]) * absorb) + \
                oracle.chain_log10_sampling_weight(ext, float(ds), float(lo), N)
            assert r["log10w"][p] == pytest.approx(want, rel=1e-6), p
