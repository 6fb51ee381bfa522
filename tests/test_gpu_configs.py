"""The benchmark configurations against the oracle at their full size (parity bars stated per test).

* 512^3 forced configuration: the exact kernel instances behind the headline number (FP64 and
  FP32-bulk x, y, z and epilogue passes at N = 512) -- one nonlinear term of each precision and one
  FP64 RK4-CN step with its energy spectrum, against the oracle's NonlinearRHSBatch / RK4Step
  (Solver/src/solver.c:505-731) and ComputeSystemMeasurables (:1142-1276).
* 256^3 decaying configuration (nu = 2e-3, k^4 start with k0 = 4, energy 0.5, seed 1, CFL 0.5):
  one FP64 step against the oracle; FP64 against FP32-bulk over 20 CFL steps with the per-output
  bars of the FP32-bulk contract (E(k) <= 1e-5 on shells holding >= 1e-10 of the peak shell).
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[configs] {msg}", flush=True)


def test_512_forced_rhs_and_step_vs_oracle(nsg, oracle):
    N, nu = 512, 1e-3
    with nsg.Solver(N, nu=nu, dt=1e-3, force_eps=0.1, force_kf=2.5) as s:
        s.init("k4", k0=2.0, energy=0.5, seed=2)
        u0 = s.get_state()
        nl64 = s.rhs(u0)
        s.step(1)
        u1 = s.get_state()
        _, ek1, _ = s.measures()
    with nsg.Solver(N, nu=nu, dt=1e-3, force_eps=0.1, force_kf=2.5, precision="fp32bulk") as s:
        nl32 = s.rhs(u0)
    _say("512^3 GPU done; oracle nonlinear term")
    want = oracle.rhs(u0, 0.1, 2.5)
    e64, e32 = rel(nl64, want), rel(nl32, want)
    _say(f"512^3 rhs rel err fp64 {e64:.2e} fp32bulk {e32:.2e}")
    assert e64 < 1e-12
    assert e32 < 2e-6
    del nl64, nl32, want
    _say("512^3 oracle RK4-CN step")
    u1o = oracle.step(u0, nu, 1e-3, 1, "rk4cn", 0.1, 2.5)
    e_inc = rel(u1 - u0, u1o - u0)
    _say(f"512^3 step increment rel err {e_inc:.2e}")
    assert e_inc < 1e-12 and rel(u1, u1o) < 1e-12
    _, eko, _ = oracle.measures(u1o, nu)
    assert rel(ek1, eko) < 1e-12


def test_256_decay_fp64_step_vs_oracle(nsg, oracle):
    N, nu = 256, 2e-3
    with nsg.Solver(N, nu=nu, dt=1.0, cfl=0.5) as s:
        s.init("k4", k0=4.0, energy=0.5, seed=1)
        u0 = s.get_state()
        s.step(1)
        u1 = s.get_state()
        _, dt, mx = s.time()
    assert abs(dt - oracle.cfl_dt(u0, 0.5)) < 1e-13 * dt
    assert abs(mx - oracle.max_speed_sum(u0)) < 1e-13 * mx
    u1o = oracle.step(u0, nu, dt)
    assert rel(u1 - u0, u1o - u0) < 1e-12 and rel(u1, u1o) < 1e-12


def test_256_decay_fp32bulk_vs_fp64_20_cfl_steps(nsg):
    """both precisions from the same start, each with its own CFL step (CFL 0.5), compared every 5
    steps: E(k) rel <= 1e-5 on resolved shells, energy rel <= 1e-8, dissipation and enstrophy rel
    <= 1e-6, state rel L2 <= 1e-6, dt rel <= 1e-6"""
    N, nu = 256, 2e-3
    kw = dict(nu=nu, dt=1.0, cfl=0.5)
    with nsg.Solver(N, **kw) as a, nsg.Solver(N, precision="fp32bulk", **kw) as b:
        a.init("k4", k0=4.0, energy=0.5, seed=1)
        b.set_state(a.get_state())
        for blk in range(4):
            a.step(5)
            b.step(5)
            ta, ea, _ = a.measures()
            tb, eb, _ = b.measures()
            sel = ea >= 1e-10 * ea.max()
            e_k = np.max(np.abs(eb[sel] - ea[sel]) / ea[sel])
            e_u = rel(b.get_state(), a.get_state())
            (tta, dta, _), (ttb, dtb, _) = a.time(), b.time()
            _say(f"256^3 decay after {5 * (blk + 1)} steps: E(k) {e_k:.2e} state {e_u:.2e} dt {abs(dtb - dta) / dta:.1e}")
            assert e_k < 1e-5
            assert abs(tb.energy - ta.energy) < 1e-8 * ta.energy
            assert abs(tb.enstrophy - ta.enstrophy) < 1e-6 * ta.enstrophy
            assert abs(tb.dissipation - ta.dissipation) < 1e-6 * ta.dissipation
            assert e_u < 1e-6
            assert abs(dtb - dta) < 1e-6 * dta and abs(ttb - tta) < 1e-6 * tta
