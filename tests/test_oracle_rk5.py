"""Oracle Dormand-Prince integrators (RK5Step / RK5DPStep, Solver/src/solver.c:217-421, step
control :134-153, tableau :34-43, tolerances Solver/src/data_types.h:123-128).

Pins, chosen so that a wrong tableau entry, weight or norm term fails one of them:
  * fifth-order self-convergence of the inviscid update (a wrong A, B or C entry drops the order);
  * the reference's own tableau (A32 = 1/2, D18) is first order, which is why it is corrected;
  * exact viscous decay of a Beltrami field under the Crank-Nicolson factor;
  * forced Beltrami flow reduces to the scalar ODE s' = eps / (2 E0 s) (u = s u0): the update and
    the embedded error norm must equal the published Dormand-Prince 5(4) pair applied to that
    scalar ODE (coefficients typed from Dormand & Prince, J. Comput. Appl. Math. 6 (1980) 19-26),
    and the adaptive controller must retrace the same accept / reject sequence;
  * the embedded error of a turbulent field scales as dt^5.
"""
import numpy as np
import pytest

from conftest import rel

# Dormand & Prince (1980), RK5(4)7M
DP_C = [0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0]
DP_A = [[],
        [1 / 5],
        [3 / 40, 9 / 40],
        [44 / 45, -56 / 15, 32 / 9],
        [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
        [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
        [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]]
DP_B = [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0]
DP_BS = [5179 / 57600, 0.0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40]


def test_published_tableau_is_consistent():
    """row sums equal the nodes; both weight sets sum to one (the checks D18 fails)"""
    for a, c in zip(DP_A[1:], DP_C[1:]):
        assert abs(sum(a) - c) < 1e-15
    assert abs(sum(DP_B) - 1) < 1e-15 and abs(sum(DP_BS) - 1) < 1e-15
    ref_row3 = 3 / 40 + 0.5
    assert abs(ref_row3 - DP_C[2]) > 0.2


def scalar_dp(s, h, f):
    K = []
    for i in range(7):
        y = s + h * sum(a * k for a, k in zip(DP_A[i], K))
        K.append(f(y))
    s5 = s + h * sum(b * k for b, k in zip(DP_B, K))
    s4 = s + h * sum(b * k for b, k in zip(DP_BS, K))
    return s5, s5 - s4


@pytest.mark.parametrize("quirk,lo,hi", [(False, 24.0, 40.0), (True, 1.6, 2.6)])
def test_rk5_order_inviscid(oracle, quirk, lo, hi):
    """Taylor-Green x3, nu = 0, T = 0.4: the error ratio for dt -> dt/2 is about 2^5 with the
    published tableau and about 2 with the reference's A32 = 1/2"""
    N, T = 12, 0.4
    u0 = oracle.ic_taylor_green(N) * 3.0

    def run(n, q):
        u = u0
        for _ in range(n):
            u, _ = oracle.step_rk5(u, 0.0, T / n, 6, quirk=q)
        return u

    ref = run(64, False)
    e1, e2 = rel(run(4, quirk), ref), rel(run(8, quirk), ref)
    assert lo < e1 / e2 < hi, (e1, e2)


def test_rk5_matches_step_scheme(oracle):
    """oracle_step's rk5cn and dp5 schemes are the six-stage update"""
    uh = oracle.ic_k4(12, 3.0, 0.5, 5)
    want, _ = oracle.step_rk5(uh, 0.01, 2e-3, 6)
    assert np.array_equal(oracle.step(uh, 0.01, 2e-3, scheme="rk5cn"), want)
    assert np.array_equal(oracle.step(uh, 0.01, 2e-3, scheme="dp5"), want)
    u7, err = oracle.step_rk5(uh, 0.01, 2e-3, 7)
    assert np.array_equal(u7, want) and err > 0


@pytest.mark.parametrize("visc_pow", [1, 2])
def test_rk5_beltrami_exact_decay(oracle, visc_pow):
    """NL = 0 for the Beltrami field: each step multiplies every mode by (2-D)/(2+D),
    D = nu lambda^(2p) dt, and the embedded error is zero (error norm 0)"""
    N, nu, dt, n = 12, 0.01, 0.01, 4
    u = u0 = oracle.ic_shapiro(N)
    for _ in range(n):
        u, err = oracle.step_rk5(u, nu, dt, 7, visc_pow=visc_pow)
        assert err < 1e-6
    D = nu * 12.0 ** visc_pow * dt
    assert rel(u, u0 * ((2 - D) / (2 + D)) ** n) < 1e-13


def _beltrami_forced(oracle, N=12, amp=1.0):
    u0 = oracle.ic_shapiro(N) * amp
    tot, _, _ = oracle.measures(u0, 0.0)
    E0 = tot[0] / (8 * np.pi ** 3)  # volume average
    return u0, E0


def _norm_of_scalar(u0, s0, s5, es):
    """the reference error norm for u = s u0, e = es u0 (per-component rms over Nx Ny Nzf)"""
    a0 = np.abs(u0)
    den = 1e-7 + np.maximum(a0 * abs(s0), a0 * abs(s5)) * 1e-7
    r = (np.abs(es * u0) / den) ** 2
    cnt = u0.shape[0] * u0.shape[1] * u0.shape[2]
    return max(np.sqrt(r[..., c].sum() / cnt) for c in range(3))


@pytest.mark.parametrize("dt", [0.02, 0.1])
def test_dp5_forced_beltrami_scalar_reduction(oracle, dt):
    """injection eps = 50 E0 makes the scalar ODE stiff enough for an O(1) error norm"""
    u0, E0 = _beltrami_forced(oracle)
    eps, kf = 50.0 * E0, 4.0
    f = lambda s: eps / (2 * E0 * s)  # noqa: E731
    s5, es = scalar_dp(1.0, dt, f)
    u5, err = oracle.step_rk5(u0, 0.0, dt, 7, force_eps=eps, force_kf=kf)
    assert rel(u5, s5 * u0) < 1e-13
    assert abs(s5 - np.sqrt(1 + eps * dt / E0)) < 0.03 * s5  # eps dt / E0 = 1 and 5: coarse steps
    want = _norm_of_scalar(u0, 1.0, s5, es)
    assert want > 1.0
    assert abs(err - want) < 1e-6 * want + 1e-9, (err, want)


def test_dp5_controller_factor(oracle):
    assert oracle.dp5_factor(1.0) == 0.8
    assert oracle.dp5_factor(1e-12) == 1.5
    assert oracle.dp5_factor(0.0) == 1.5
    assert oracle.dp5_factor(1e12) == 0.01
    assert abs(oracle.dp5_factor(32.0) - 0.4) < 1e-15
    assert np.isnan(oracle.dp5_factor(float("nan")))


def test_dp5_solve_retraces_scalar_controller(oracle):
    """adaptive run of forced Beltrami flow from an oversized dt: same accepted times, dt and
    rejection count as the scalar Dormand-Prince pair under the same controller"""
    u0, E0 = _beltrami_forced(oracle)
    eps, kf, nsteps = 50.0 * E0, 4.0, 6
    f = lambda s: eps / (2 * E0 * s)  # noqa: E731
    s, t, dt, rej = 1.0, 0.0, 1.0, 0
    for _ in range(nsteps):
        while True:
            s5, es = scalar_dp(s, dt, f)
            err = _norm_of_scalar(u0, s, s5, es)
            h = dt
            dt = h * min(1.5, max(0.01, 0.8 * err ** -0.2))
            if err <= 1.0:
                s, t = s5, t + h
                break
            rej += 1
    u, t_o, dt_o, rej_o, _, rc = oracle.dp5_solve(u0, 0.0, 1.0, nsteps, force_eps=eps, force_kf=kf)
    assert rc == 0
    assert rej_o == rej and rej >= 1
    assert abs(t_o - t) < 1e-9 * t and abs(dt_o - dt) < 1e-6 * dt
    assert rel(u, s * u0) < 1e-10
    assert abs(s - np.sqrt(1 + eps * t / E0)) < 1e-5


def test_dp5_error_scales_as_dt5(oracle):
    """k^4 turbulence at N = 16, nu = 0: err(dt) / err(dt/2) is about 2^5"""
    uh = oracle.ic_k4(16, 3.0, 0.5, 3)
    dt = 4e-3
    e1 = oracle.step_rk5(uh, 0.0, dt, 7)[1]
    e2 = oracle.step_rk5(uh, 0.0, dt / 2, 7)[1]
    assert 26.0 < e1 / e2 < 38.0, (e1, e2)


def test_dp5_solve_gives_up(oracle):
    """a non-finite state ends the run with rc = -1 instead of looping"""
    uh = oracle.ic_k4(12, 3.0, 0.5, 1)
    uh[1, 1, 1, 0] = np.nan
    *_, rc = oracle.dp5_solve(uh, 0.01, 1e-3, 1)
    assert rc == -1
