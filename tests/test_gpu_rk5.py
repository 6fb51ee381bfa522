"""Dormand-Prince integrators on the GPU (NSG_SCHEME_RK5CN = the reference's -D__RK5 step,
NSG_SCHEME_DP5 = its adaptive -D__DPRK5 loop, Solver/src/solver.c:134-153, 217-421) through the
C ABI, against the CPU oracle (oracle_step_rk5 / oracle_dp5_solve) on identical seeded states.

Bars: FP64 state rel L2 <= 1e-12 per fixed step and <= 1e-11 after adaptive runs; the adaptive
runs must take the same accept / reject decisions (same rejection count, t and next dt to 1e-9
relative) and report the same error norm (1e-8 relative: a sum of squared ratios over the
whole spectrum, reduced in a different order).  FP32-bulk RK5-CN: one fixed step rel <= 1e-8 vs
the FP64 oracle (the adaptive scheme is FP64 only).
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def k4(oracle, N, seed=None):
    return oracle.ic_k4(N, min(2.0, N / 8.0), 0.5, seed if seed is not None else 1000 + N)


CASES = [(1, 0.0), (2, 0.0), (1, 0.1)]


@pytest.mark.parametrize("N", [8, 12, 16, 24, 32, 48, 64])
@pytest.mark.parametrize("p,eps", CASES)
def test_rk5cn_steps_fp64(nsg, oracle, N, p, eps):
    uh = k4(oracle, N)
    nu = 0.01 if p == 1 else 1e-5
    dt = 2e-3
    with nsg.Solver(N, nu=nu, dt=dt, visc_pow=p, scheme="rk5cn", force_eps=eps) as s:
        s.set_state(uh)
        s.step(2)
        got = s.get_state()
        t, dtn, _ = s.time()
    want = oracle.step(uh, nu, dt, p, "rk5cn", eps, 2.5, nsteps=2)
    assert rel(got - uh, want - uh) < 1e-12
    assert rel(got, want) < 1e-12
    assert abs(t - 2 * dt) < 1e-15 and dtn == dt


@pytest.mark.parametrize("N", [32, 64])
def test_rk5cn_fp32bulk(nsg, oracle, N):
    uh = k4(oracle, N)
    with nsg.Solver(N, nu=0.01, dt=2e-3, scheme="rk5cn", precision="fp32bulk") as s:
        s.set_state(uh)
        s.step(1)
        got = s.get_state()
    assert rel(got, oracle.step(uh, 0.01, 2e-3, 1, "rk5cn")) < 1e-8


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32bulk", 1e-6)])
def test_rk5cn_beltrami_decay(nsg, precision, tol):
    """NL = 0: every step multiplies the field by (2-D)/(2+D), D = nu lambda^2 dt"""
    N, nu, dt = 32, 0.01, 0.01
    with nsg.Solver(N, nu=nu, dt=dt, scheme="rk5cn", precision=precision) as s:
        s.init("shapiro")
        u0 = s.get_state()
        s.step(5)
        D = nu * 12 * dt
        assert rel(s.get_state(), u0 * ((2 - D) / (2 + D)) ** 5) < tol


@pytest.mark.parametrize("N,nu,dt0,eps", [(16, 0.01, 0.3, 0.0), (32, 0.01, 0.3, 0.0), (24, 0.01, 0.05, 0.1),
                                          (48, 2e-3, 0.02, 0.0)])
def test_dp5_adaptive_vs_oracle(nsg, oracle, N, nu, dt0, eps):
    uh = k4(oracle, N)
    nsteps = 4
    with nsg.Solver(N, nu=nu, dt=dt0, scheme="dp5", force_eps=eps) as s:
        s.set_state(uh)
        s.step(nsteps)
        got = s.get_state()
        t, dt, _ = s.time()
        acc, rej, last = s.dp_stats()
    want, t_o, dt_o, rej_o, last_o, rc = oracle.dp5_solve(uh, nu, dt0, nsteps, force_eps=eps, force_kf=2.5)
    assert rc == 0
    assert acc == nsteps and rej == rej_o, (rej, rej_o)
    # dt = dt err^-1/5 carries the error norm's reduction-order round-off into t and dt
    assert abs(t - t_o) < 1e-9 * t_o and abs(dt - dt_o) < 1e-9 * dt_o
    assert abs(last - last_o) < 1e-8 * last_o
    assert rel(got, want) < 1e-11


def test_dp5_rejects_oversized_first_try(nsg, oracle):
    """dt0 = 0.3 at N = 16: the first try is rejected, dt shrinks, the state does not move on a
    rejection and time advances by the accepted dt only"""
    N = 16
    uh = k4(oracle, N)
    with nsg.Solver(N, nu=0.01, dt=0.3, scheme="dp5") as s:
        s.set_state(uh)
        s.step(1)
        acc, rej, _ = s.dp_stats()
        t, dt, _ = s.time()
    want, t_o, dt_o, rej_o, _, _ = oracle.dp5_solve(uh, 0.01, 0.3, 1)
    assert rej_o >= 1 and rej == rej_o and acc == 1
    assert abs(t - t_o) < 1e-9 * t_o and t < 0.3


def test_dp5_nonfinite_state_fails(nsg, oracle):
    """a NaN in the state gives a NaN error norm: NSG_ERR_NONFINITE, not an endless retry"""
    N = 16
    uh = k4(oracle, N)
    uh[1, 1, 1, 0] = np.nan
    with nsg.Solver(N, nu=0.01, dt=1e-3, scheme="dp5") as s:
        s.set_state(uh)
        with pytest.raises(nsg.NSGError) as e:
            s.step(1)
        assert e.value.code == -4


def test_dp5_stats_only_for_dp5(nsg):
    with nsg.Solver(16, scheme="rk5cn") as s:
        with pytest.raises(nsg.NSGError):
            s.dp_stats()


def test_dp5_autotune_restores_state(nsg, oracle):
    N = 48
    uh = k4(oracle, N)
    with nsg.Solver(N, nu=2e-3, dt=0.02, scheme="dp5") as s:
        s.set_state(uh)
        s.step(1)
        before, tb = s.get_state(), s.time()
        chosen, _ = s.autotune(iters=2, use_cache=False)
        assert np.array_equal(s.get_state(), before) and s.time() == tb
        s.step(2)
        got, tg = s.get_state(), s.time()
    with nsg.Solver(N, nu=2e-3, dt=0.02, scheme="dp5") as s:
        s.set_state(uh)
        s.step(1)
        s.set_schedule(chosen)
        s.step(2)
        assert np.array_equal(s.get_state(), got) and s.time() == tg


def test_rk5cn_ensemble_equals_single_box(nsg):
    """RK5-CN boxes inside a batch / graph are bitwise the same boxes stepped alone, adaptive CFL
    dt included"""
    boxes = [dict(N=32, nu=2e-3, dt=1e-3, scheme="rk5cn", precision="fp32bulk", cfl=0.5),
             dict(N=32, nu=4e-3, dt=1e-3, scheme="rk5cn", precision="fp32bulk", cfl=0.5),
             dict(N=24, nu=1e-2, dt=2e-3, scheme="rk5cn", force_eps=0.1, force_kf=2.5)]
    with nsg.Ensemble(boxes) as e:
        assert e.groups == 2
        for b in range(len(boxes)):
            e.init(b, "k4", k0=2.0 + b, energy=0.5, seed=77 + b)
        e.step(3)
        got = [(e.get_state(b), e.time(b)) for b in range(len(boxes))]
    for b, p in enumerate(boxes):
        with nsg.Solver(**p) as s:
            s.init("k4", k0=2.0 + b, energy=0.5, seed=77 + b)
            s.step(3)
            assert np.array_equal(got[b][0], s.get_state()), f"box {b}"
            assert got[b][1] == s.time()


def test_dp5_with_zero_closure_is_plain_dp5(nsg, oracle):
    """a closure network with zero weights adds a zero term: bitwise the plain adaptive run"""
    N = 32
    uh = k4(oracle, N)
    with nsg.Solver(N, nu=0.01, dt=0.05, scheme="dp5") as s:
        s.set_state(uh)
        s.step(2)
        want = s.get_state()
    with nsg.Network(N, seed=1) as net, nsg.Solver(N, nu=0.01, dt=0.05, scheme="dp5") as s:
        net.set_params(np.zeros(net.param_count, dtype=np.float32))
        s.attach_closure(net)
        s.set_state(uh)
        s.step(2)
        assert np.array_equal(s.get_state(), want)


def test_dp5_profiling_leaves_state_and_time(nsg, oracle):
    """pass profiling runs tries without the accept step: state, time and dt stay as they were"""
    N = 32
    with nsg.Solver(N, nu=0.01, dt=0.02, scheme="dp5") as s:
        s.set_state(k4(oracle, N))
        s.step(1)
        before, tb = s.get_state(), s.time()
        ms, by = s.profile(2)
        assert all(m > 0 for m in ms[:5]) and all(b > 0 for b in by)
        assert np.array_equal(s.get_state(), before) and s.time() == tb
