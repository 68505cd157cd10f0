"""The injection oracle (oracle/inject_oracle.py) pinned on the CPU: the
generator against the published SplitMix64 sequence and a scalar restatement
in Python integers, the distributions against their closed forms (the cosine
profile of generate_pairs, uniform y, normal velocities with variance |q/m| T), the
pair selection against a brute-force sequential loop (the reference's
attempt loop, utils.cpp:636-683), and the counts of main.cpp:235-238 and
utils.cpp:619-635."""
import math

import numpy as np
from scipy import stats

import inject_oracle as IO

E, ME, MI = 1.602176634e-19, 9.1093837015e-31, 40 * 1.66053906660e-27
MASK = (1 << 64) - 1


def _mix_int(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def _urand_int(seed, step, stream, k, j):
    z = _mix_int((seed + IO.GOLDEN * (step + 1)) & MASK)
    z = _mix_int(z ^ (stream << 60) ^ (k << 8) ^ j)
    return ((z >> 12) + 0.5) / 2.0 ** 52


def test_splitmix64_published_sequence():
    # first outputs of SplitMix64 seeded with 1234567 (the generator's reference test vector)
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert [int(v) for v in IO.splitmix64(1234567, 5)] == want


def test_urand_matches_integer_restatement_and_range():
    k = np.arange(2000, dtype=np.uint64)
    for seed, step, stream, j in [(1, 0, 0, 0), (7, 3, 1, 5), (2 ** 63 + 5, 1000, 1, 9)]:
        u = IO.urand(seed, step, stream, k, j)
        ref = np.array([_urand_int(seed, step, stream, int(i), j) for i in k[:200]])
        assert np.array_equal(u[:200], ref)
        assert (u > 0).all() and (u < 1).all()
    u = IO.urand(3, 0, 0, np.arange(200000, dtype=np.uint64), 0)
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    # draws j and j + 1 are independent
    v = IO.urand(3, 0, 0, np.arange(200000, dtype=np.uint64), 1)
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.01


def test_maxwellian_components_are_normal_with_variance_qm_T():
    k = np.arange(100000, dtype=np.uint64)
    T = 2.5
    sig = math.sqrt(E / ME * T)
    vx, vy, vz = IO.maxwell(9, 2, 0, k, 1, sig)
    for v in (vx, vy, vz):
        assert stats.kstest(v / sig, "norm").pvalue > 1e-3
    c = np.corrcoef([vx, vy, vz])
    assert np.abs(c - np.eye(3)).max() < 0.02
    # the reference's v * sqrt(-ln U) cos(2 pi U'), v = sqrt(2 |q/m| T), has the same variance
    assert math.isclose(math.sqrt(2 * E / ME * T) ** 2 / 2, sig ** 2)


def test_pair_x_follows_the_cosine_profile():
    """x = xm + R asin(2u - 1) has CDF (1 + sin((x - xm)/R)) / 2: density
    cos((x - xm)/R) / (2R), peaked at xm and vanishing at the zone ends."""
    u = IO.urand(4, 0, 1, np.arange(200000, dtype=np.uint64), 0)
    x = IO.pair_x(u)
    R = 0.0075 / IO.PI_REF     # half-width of the support is R * pi / 2
    assert x.min() >= 0.00625 - R * math.pi / 2 and x.max() <= 0.00625 + R * math.pi / 2
    cdf = lambda t: (np.sin(np.clip((t - 0.00625) / R, -math.pi / 2, math.pi / 2)) + 1) / 2
    assert stats.kstest(x, cdf).pvalue > 1e-3
    # symmetric about xm, density highest in the middle
    assert abs(np.median(x) - 0.00625) < 2e-5
    h, _ = np.histogram(x, bins=10)
    assert h[5] > 4 * h[0] and h[4] > 4 * h[-1]


def test_pair_selection_is_the_sequential_loop():
    """Brute force: walk attempts in order with the scalar generator, keep
    0 < x < width, stop at target (utils.cpp:636-683)."""
    seed, step, width, height = 5, 3, 0.0064, 0.0064
    target, attempts = 170, 1700
    e, i = IO.pairs(seed, step, target, attempts, width, height, 1.0, 2.0)
    kept = []
    for a in range(attempts):
        if len(kept) == target:
            break
        x = 0.00625 + math.asin(2.0 * _urand_int(seed, step, 1, a, 0) - 1.0) * 0.0075 / IO.PI_REF
        if 0.0 < x < width:
            kept.append((a, x))
    assert len(kept) == target == e[0].size == i[0].size
    np.testing.assert_allclose(e[0], [x for _, x in kept], rtol=1e-15)
    ys = [height * _urand_int(seed, step, 1, a, 1) for a, _ in kept]
    np.testing.assert_allclose(e[1], ys, rtol=1e-15)
    assert np.array_equal(e[0], i[0]) and np.array_equal(e[1], i[1])
    # ions from draws 6-9, electrons from draws 2-5: independent velocities
    assert not np.allclose(e[2], i[2] / 2.0)
    # too few accepted attempts: fewer pairs than the target
    e2, _ = IO.pairs(seed, step, 10 ** 6, 500, width, height, 1.0, 1.0)
    assert 0 < e2[0].size < 500


def test_counts():
    assert IO.reinject_count(10, 3) == 7
    assert IO.reinject_count(3, 10) == 0
    assert IO.reinject_count(0, 0) == 0
    assert IO.pair_target(0.0064, 1.77e-11, x11=0.01, x22=0.01) == (0, 0)
    # linear in height and dt; target is the ceiling; attempts max(10 target, 100)
    r1 = IO.pair_rate(0.0064, 1.77e-11)
    assert math.isclose(IO.pair_rate(0.0128, 1.77e-11), 2 * r1, rel_tol=1e-15)
    assert math.isclose(IO.pair_rate(0.0064, 3.54e-11), 2 * r1, rel_tol=1e-15)
    t, a = IO.pair_target(0.0064, 1.77e-11)
    assert t == math.ceil(r1) and t - 1 < r1 <= t and a == max(10 * t, 100)
    t_small, a_small = IO.pair_target(1e-6, 1e-12)
    assert t_small == 1 and a_small == 100
    # the rate equals the ionisation profile's integral: mean over the zone of
    # ionimax * height * dt / 1.67e6 with the mean of sin over half a period (2/pi)
    r_direct = 0.0064 * 1.77e-11 * 5.23e23 / 1.67e6 * 0.0075 * 2.0 / IO.PI_REF
    assert math.isclose(r1, r_direct, rel_tol=1e-14)


def test_injection_bundle_order_and_clamps():
    kw = dict(width=0.0064, height=0.0064, dt=1.77e-11, qm_e=-E / ME, qm_i=E / MI)
    e, i = IO.injection(1, 0, 50, 20, x_inject=0.003, **kw)
    t, _ = IO.pair_target(0.0064, 1.77e-11)
    assert e[0].size == 30 + t and i[0].size == t
    assert (e[0][:30] == 0.003).all()
    # plane outside the grid: nothing re-injected (the reference drops them)
    e, _ = IO.injection(1, 0, 50, 20, x_inject=0.024, **kw)
    assert e[0].size == t
    e, _ = IO.injection(1, 0, 50, 20, x_inject=0.003, max_reinject=7, **kw)
    assert e[0].size == 7 + t
    e, i = IO.injection(1, 0, 50, 20, x_inject=0.003, pairs_on=False, **kw)
    assert e[0].size == 30 and i[0].size == 0
