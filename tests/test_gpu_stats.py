"""Velocity-increment statistics on the GPU (nsg_incr_*, the reference's post-processing
PostProcessing/src/stats.c) against the CPU oracle (oracle/stats_oracle.c) on identical fields.

Bars: counts, min, max and the bin ranges bitwise equal to the oracle's (the bin search evaluates
the ranges exactly as the host does); mean within 1e-12 of the standard deviation; sd rel <= 1e-10;
skew and excess kurtosis within 1e-9 relative or 1e-10 absolute (they are near 0 for these
near-Gaussian fields); the differences come from parallel moment merging vs the reference's
sequential update.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def field(oracle, N, seed, amp=1.0):
    return amp * oracle.realspace(oracle.ic_k4(N, min(3.0, N / 6.0), 0.5, seed), 0)


def check_stats(st, acc, stats_o):
    for q in range(9):
        if acc[q, 0] == 0:
            assert np.all(st[q] == 0)
            continue
        assert st[q, 0] == stats_o[q, 0] and st[q, 1] == stats_o[q, 1]
        assert abs(st[q, 2] - stats_o[q, 2]) <= 1e-12 * stats_o[q, 3]
        assert abs(st[q, 3] - stats_o[q, 3]) <= 1e-10 * stats_o[q, 3]
        # skew and excess kurtosis of near-Gaussian fields are close to 0: absolute bar
        assert np.allclose(st[q, 4:], stats_o[q, 4:], rtol=1e-9, atol=1e-10), (q, st[q], stats_o[q])


@pytest.mark.parametrize("N", [8, 12, 16, 24, 32, 48, 64])
def test_increment_stats_vs_oracle(nsg, oracle, N):
    snaps = [field(oracle, N, 100 + N), field(oracle, N, 200 + N, 1.5)]
    acc = None
    for u in snaps:
        acc = oracle.incr_moments(u, acc)
    ranges_o = oracle.incr_ranges(acc)
    counts_o = None
    for u in snaps:
        counts_o, out = oracle.incr_histogram(u, ranges_o, counts_o)
        assert out == 0
    with nsg.IncrementStats(N) as s:
        assert list(s.increments) == list(oracle.incr_list(N))
        for u in snaps:
            s.precompute(u)
        s.set_ranges()
        for u in snaps:
            s.histogram(u)
        st, rg, ct = s.results()
    check_stats(st, acc, oracle.incr_stats(acc))
    assert np.array_equal(rg, ranges_o)
    assert np.array_equal(ct, counts_o)
    assert np.array_equal(ct.sum(axis=1), acc[:, 0])


def test_increment_stats_from_solver_state(nsg, oracle):
    """the device-resident state's real-space velocity is the same field nsg_realspace returns;
    scale = N^3 gives the post-processing's unnormalised field (exact for N = 32)"""
    N = 32
    with nsg.Solver(N) as sol, nsg.IncrementStats(N) as a, nsg.IncrementStats(N) as b, \
            nsg.IncrementStats(N) as c:
        sol.init("k4", k0=3.0, energy=0.5, seed=9)
        u = sol.realspace(0)
        a.precompute(sol)
        b.precompute(u)
        c.precompute(sol, scale=N ** 3)
        for s, snap, sc in ((a, sol, 1.0), (b, u, 1.0), (c, sol, N ** 3)):
            s.set_ranges()
            s.histogram(snap, sc)
        sa, ra, ca = a.results()
        sb, rb, cb = b.results()
        sc_, _, cc = c.results()
    assert np.array_equal(sa, sb) and np.array_equal(ra, rb) and np.array_equal(ca, cb)
    assert np.allclose(sc_[:, :4], sa[:, :4] * N ** 3, rtol=1e-12, atol=0)
    assert np.array_equal(cc.sum(axis=1), ca.sum(axis=1))


def test_linear_ramp_closed_form(nsg, oracle):
    """every increment a r sits on the middle bin edge of its histogram: the exact-range path of
    the bin search must pick the oracle's bin"""
    N, a = 16, 0.5
    i = np.arange(N, dtype=np.float64)
    u = np.zeros((N, N, N, 3))
    u[..., 0] = a * i[:, None, None]
    u[..., 1] = a * i[None, :, None]
    u[..., 2] = a * i[None, None, :]
    with nsg.IncrementStats(N) as s:
        s.precompute(u)
        s.set_ranges()
        s.histogram(u)
        st, rg, ct = s.results()
    want, out = oracle.incr_histogram(u, rg)
    assert out == 0 and np.array_equal(ct, want)
    for q, r in enumerate([1, 2, 4, 8, 16, 32, 64, 128, 8]):
        if r >= N:
            assert np.all(st[q] == 0) and ct[q].sum() == 0
            continue
        assert st[q, 0] == st[q, 1] == st[q, 2] == a * r and st[q, 3] == 0.0
        assert ct[q].sum() == 3 * N * N * (N - r) and np.count_nonzero(ct[q]) == 1


def test_outside_ranges_and_order_errors(nsg, oracle):
    N = 12
    u = field(oracle, N, 3)
    with nsg.IncrementStats(N) as s:
        with pytest.raises(nsg.NSGError) as e:
            s.histogram(u)
        assert e.value.code == -1 and "set_ranges" in str(e.value)
        s.precompute(u)
        s.set_ranges()
        with pytest.raises(nsg.NSGError) as e:
            s.histogram(5.0 * u)
        assert e.value.code == -1 and "outside the bin ranges" in str(e.value)
        _, rg, ct = s.results()
    want, out = oracle.incr_histogram(5.0 * u, rg)
    assert out > 0 and np.array_equal(ct, want)


def test_nonfinite_field(nsg, oracle):
    N = 12
    u = field(oracle, N, 4)
    u[3, 4, 5, 1] = np.nan
    with nsg.IncrementStats(N) as s:
        s.precompute(u)
        with pytest.raises(nsg.NSGError) as e:
            s.set_ranges()
        assert e.value.code == -4


def test_increment_stats_256_vs_numpy(nsg, oracle):
    """full-size properties at 256^3: every increment counted once (3 N^2 (N - r)), the moments of
    three separations against numpy's two-pass values, and a bitwise repeat"""
    N = 256
    with nsg.Solver(N) as sol:
        sol.init("k4", k0=4.0, energy=0.5, seed=21)
        u = sol.realspace(0)
    res = []
    for _ in range(2):
        with nsg.IncrementStats(N) as s:
            s.precompute(u)
            s.set_ranges()
            s.histogram(u)
            res.append(s.results())
    (st, rg, ct), (st2, rg2, ct2) = res
    assert np.array_equal(st, st2) and np.array_equal(ct, ct2)
    for q, r in enumerate([1, 2, 4, 8, 16, 32, 64, 128, 128]):
        assert ct[q].sum() == 3 * N * N * (N - r)
    for q, r in ((0, 1), (4, 16), (8, 128)):
        x = np.concatenate([(u[r:, :, :, 0] - u[:-r, :, :, 0]).ravel(), (u[:, r:, :, 1] - u[:, :-r, :, 1]).ravel(),
                            (u[:, :, r:, 2] - u[:, :, :-r, 2]).ravel()])
        m, sd = x.mean(), x.std(ddof=1)
        d = x - m
        assert st[q, 0] == x.min() and st[q, 1] == x.max()
        assert abs(st[q, 2] - m) < 1e-12 * sd and rel(st[q, 3], sd) < 1e-10
        assert abs(st[q, 4] - np.mean(d ** 3) / sd ** 3) < 1e-9
        assert abs(st[q, 5] - (np.mean(d ** 4) / sd ** 4 - 3.0)) < 1e-9
