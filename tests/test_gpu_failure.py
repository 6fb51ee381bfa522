"""Failure detection through the C ABI: non-finite states, the blow-up guard, and restart from a
saved (u_hat, t, dt) continuing bitwise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_nonfinite_state_is_reported(nsg, oracle):
    N = 32
    uh = oracle.ic_k4(N, 2.0, 0.5, 1)
    uh[1, 2, 3, 0] = np.nan
    with nsg.Solver(N, nu=0.01) as s:
        s.set_state(uh)
        with pytest.raises(nsg.NSGError) as e:
            s.measures()
        assert e.value.code == -4
        with pytest.raises(nsg.NSGError) as e:
            s.step(1)
        assert e.value.code == -4


def test_blowup_guard(nsg, oracle):
    N = 32
    with nsg.Solver(N, nu=0.01, max_speed=1e-3) as s:
        s.init("k4", k0=2.0, energy=0.5, seed=1)
        with pytest.raises(nsg.NSGError) as e:
            s.step(1)
        assert e.value.code == -5 and "exceeds" in str(e.value)


def test_restart_continues_bitwise(nsg):
    N = 64
    kw = dict(nu=1e-3, dt=1e-3, cfl=0.5, force_eps=0.1, scheme="ifrk4")
    with nsg.Solver(N, **kw) as a:
        a.init("k4", k0=2.0, energy=0.5, seed=4)
        a.step(4)
        ref = a.get_state()
        tref = a.time()
    with nsg.Solver(N, **kw) as b:
        b.init("k4", k0=2.0, energy=0.5, seed=4)
        b.step(2)
        snap, (t, dt, _) = b.get_state(), b.time()
    with nsg.Solver(N, **kw) as c:
        c.set_state(snap)
        c.set_time(t, dt)
        c.step(2)
        assert np.array_equal(c.get_state(), ref)
        assert c.time()[:2] == tref[:2]
