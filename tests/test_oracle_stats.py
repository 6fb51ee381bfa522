"""Oracle increment statistics (PostProcessing/src/stats.c: Precompute :31-150, ComputeStats
:153-250, WriteStatsData :316-371) against independent numpy computations: the increments are
gathered by array slicing (non-periodic: i + r < N), the moments by two-pass formulas, the bins by
searchsorted on the same edges.  GSL's definitions (sd with n - 1; skew and kurtosis with /n
moments over that sd) are what numpy is asked for here."""
import numpy as np
import pytest

NI, NB = 9, 1000


def increments(u, r):
    """all longitudinal increments at separation r, in no particular order"""
    N = u.shape[0]
    if r >= N:
        return np.zeros(0)
    return np.concatenate([(u[r:, :, :, 0] - u[:-r, :, :, 0]).ravel(), (u[:, r:, :, 1] - u[:, :-r, :, 1]).ravel(),
                           (u[:, :, r:, 2] - u[:, :, :-r, 2]).ravel()])


def numpy_stats(x):
    if x.size == 0:
        return np.zeros(6)
    m = x.mean()
    sd = x.std(ddof=1) if x.size > 1 else 0.0
    d = x - m
    return np.array([x.min(), x.max(), m, sd, np.mean(d ** 3) / sd ** 3, np.mean(d ** 4) / sd ** 4 - 3.0])


def field(oracle, N, seed):
    return oracle.realspace(oracle.ic_k4(N, 3.0, 0.5, seed), 0)


def test_increment_list(oracle):
    assert list(oracle.incr_list(64)) == [1, 2, 4, 8, 16, 32, 64, 128, 32]


@pytest.mark.parametrize("N", [12, 16, 24])
def test_moments_vs_numpy(oracle, N):
    u = field(oracle, N, 7 + N)
    acc = oracle.incr_moments(u)
    st = oracle.incr_stats(acc)
    for q, r in enumerate(oracle.incr_list(N)):
        x = increments(u, int(r))
        assert acc[q, 0] == x.size
        want = numpy_stats(x)
        if x.size == 0:
            assert np.all(st[q] == 0)
            continue
        assert st[q, 0] == want[0] and st[q, 1] == want[1]
        assert abs(st[q, 2] - want[2]) < 1e-12 * want[3]
        assert np.allclose(st[q, 3:], want[3:], rtol=1e-10, atol=0)


def test_moments_accumulate_over_snapshots(oracle):
    N = 16
    u1, u2 = field(oracle, N, 1), 2.0 * field(oracle, N, 2)
    acc = oracle.incr_moments(u1)
    acc = oracle.incr_moments(u2, acc)
    st = oracle.incr_stats(acc)
    for q, r in enumerate(oracle.incr_list(N)[:4]):
        x = np.concatenate([increments(u1, int(r)), increments(u2, int(r))])
        assert np.allclose(st[q], numpy_stats(x), rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("N", [12, 16])
def test_histogram_vs_searchsorted(oracle, N):
    u = field(oracle, N, 30 + N)
    acc = oracle.incr_moments(u)
    ranges = oracle.incr_ranges(acc)
    for q in range(NI):
        lo, hi = acc[q, 1] - 0.05, acc[q, 2] + 0.05
        b = np.arange(NB + 1)
        assert np.array_equal(ranges[q], ((NB - b) / NB) * lo + (b / NB) * hi)
        assert np.all(np.diff(ranges[q]) > 0)
    counts, outside = oracle.incr_histogram(u, ranges)
    assert outside == 0
    for q, r in enumerate(oracle.incr_list(N)):
        x = increments(u, int(r))
        want = np.bincount(np.searchsorted(ranges[q], x, side="right") - 1, minlength=NB)[:NB]
        assert np.array_equal(counts[q], want), q
        assert counts[q].sum() == x.size == acc[q, 0]


def test_linear_ramp_closed_form(oracle):
    """u_c = a x_c (index units): every increment at separation r equals a r exactly; sd = 0, so
    GSL's skew and kurtosis are 0/0"""
    N, a = 16, 0.5
    i = np.arange(N, dtype=np.float64)
    u = np.zeros((N, N, N, 3))
    u[..., 0] = a * i[:, None, None]
    u[..., 1] = a * i[None, :, None]
    u[..., 2] = a * i[None, None, :]
    acc = oracle.incr_moments(u)
    st = oracle.incr_stats(acc)
    ranges = oracle.incr_ranges(acc)
    counts, outside = oracle.incr_histogram(u, ranges)
    assert outside == 0
    for q, r in enumerate(oracle.incr_list(N)):
        if r >= N:
            assert acc[q, 0] == 0 and np.all(st[q] == 0) and counts[q].sum() == 0
            continue
        assert acc[q, 0] == 3 * N * N * (N - r)
        assert st[q, 0] == st[q, 1] == a * r and st[q, 2] == a * r and st[q, 3] == 0.0
        assert np.isnan(st[q, 4]) and np.isnan(st[q, 5])
        nz = np.nonzero(counts[q])[0]
        assert nz.size == 1 and nz[0] in (NB // 2 - 1, NB // 2) and counts[q, nz[0]] == acc[q, 0]


def test_values_outside_the_ranges_are_counted(oracle):
    """ranges from one snapshot, histogram of a field five times larger: the reference would exit;
    the oracle counts the escaping increments (same number numpy finds)"""
    N = 12
    u = field(oracle, N, 3)
    ranges = oracle.incr_ranges(oracle.incr_moments(u))
    counts, outside = oracle.incr_histogram(5.0 * u, ranges)
    want = sum(int(np.count_nonzero((increments(5.0 * u, int(r)) < ranges[q, 0]) |
                                    (increments(5.0 * u, int(r)) >= ranges[q, -1])))
               for q, r in enumerate(oracle.incr_list(N)))
    assert outside == want > 0
    assert counts.sum() + outside == sum(increments(u, int(r)).size for r in oracle.incr_list(N))
