"""Full-size checks at 256^3 and 512^3: parity at 256^3 against the oracle, size-independent
properties at 512^3 (divergence, dealias support, Hermitian plane, energy conservation of the
nonlinear term, exact Beltrami decay, FP32-bulk error from identical states, determinism)."""
import numpy as np
import pytest

from conftest import rel, wavenumbers

pytestmark = pytest.mark.gpu


def test_rhs_parity_256(nsg, oracle):
    N = 256
    uh = oracle.ic_k4(N, 2.0, 0.5, 5)
    with nsg.Solver(N) as s:
        assert rel(s.rhs(uh), oracle.rhs(uh)) < 1e-12


def test_shapiro_exact_decay_256(nsg):
    N, nu, dt = 256, 1e-3, 1e-3
    with nsg.Solver(N, nu=nu, dt=dt) as s:
        s.init("shapiro")
        u0 = s.get_state()
        s.step(3)
        D = nu * 12 * dt
        assert rel(s.get_state(), u0 * ((2 - D) / (2 + D)) ** 3) < 1e-12


@pytest.mark.parametrize("scheme,precision", [("ifrk4", "fp64"), ("ifrk4", "fp32bulk"), ("rk4cn", "fp32bulk")])
def test_shapiro_exact_decay_512(nsg, scheme, precision):
    """the 512^3 kernels (pass 5 with its reads pipelined through LDS, the FP32 z pass pairing the
    P_z of two lines) on the corrected Shapiro field, which is Beltrami: the nonlinear term is zero,
    so three steps multiply the state by exactly exp(-nu lambda^2 dt)^3 (IF-RK4) or
    ((2 - D)/(2 + D))^3 (RK4-CN), lambda^2 = 12, D = nu lambda^2 dt.  FP64 to 1e-12; FP32-bulk to
    1e-6 (its transforms' round-off leaves a nonlinear term of that relative size)"""
    N, nu, dt = 512, 1e-3, 1e-3
    with nsg.Solver(N, nu=nu, dt=dt, scheme=scheme, precision=precision) as s:
        s.init("shapiro")
        u0 = s.get_state()
        s.step(3)
        got = s.get_state()
    D = nu * 12 * dt
    g = np.exp(-D) if scheme == "ifrk4" else (2 - D) / (2 + D)
    assert rel(got, u0 * g ** 3) < (1e-12 if precision == "fp64" else 1e-6)


def _half_weights(N):
    KX, KY, KZ = wavenumbers(N)
    return np.where((KZ == 0) | (KZ == N // 2), 1.0, 2.0), (KX, KY, KZ)


def test_rhs_properties_512(nsg):
    N = 512
    with nsg.Solver(N, precision="fp64") as s:
        s.init("k4", k0=2.0, energy=0.5, seed=2)
        uh = s.get_state()
        nl = s.rhs(uh)
    w, (KX, KY, KZ) = _half_weights(N)
    scale = np.max(np.abs(nl))
    div = KX * nl[..., 0] + KY * nl[..., 1] + KZ * nl[..., 2]
    assert np.max(np.abs(div)) < 1e-12 * scale * N
    k2 = KX * KX + KY * KY + KZ * KZ
    assert not np.any(nl[k2 > (N // 3) ** 2])
    p = nl[:, :, 0, 0]
    m = np.conj(p[(-np.arange(N)) % N][:, (-np.arange(N)) % N])
    assert np.max(np.abs(p - m)) < 1e-12 * scale
    tr = np.sum(w[..., None] * np.real(np.conj(uh) * nl))
    assert abs(tr) < 1e-11 * np.sum(w[..., None] * np.abs(uh) * np.abs(nl))


def test_fp32bulk_vs_fp64_512(nsg):
    """one step from an identical forced-turbulence state: E(k) within 1e-5 of FP64 on every shell
    holding at least 1e-10 of the peak shell's energy (below that FP32 round-off is a noise floor)"""
    N = 512
    kw = dict(nu=1e-3, dt=1e-3, force_eps=0.1)
    with nsg.Solver(N, **kw) as a, nsg.Solver(N, precision="fp32bulk", **kw) as b:
        a.init("k4", k0=2.0, energy=0.5, seed=2)
        b.set_state(a.get_state())
        a.step(1)
        b.step(1)
        ta, ea, _ = a.measures()
        tb, eb, _ = b.measures()
    sel = ea >= 1e-10 * ea.max()
    assert np.max(np.abs(eb[sel] - ea[sel]) / ea[sel]) < 1e-5
    assert abs(tb.energy - ta.energy) < 1e-7 * ta.energy
    assert abs(tb.dissipation - ta.dissipation) < 1e-6 * ta.dissipation


def test_fp32bulk_vs_fp64_1024(nsg):
    """1024^3, where pass 1 takes 64-byte row segments in pairs of blocks and the two precisions use
    different x-pass layouts: one step from the same k^4 state agrees in E(k) to 1e-5 on the shells
    holding at least 1e-10 of the peak shell's energy, and in energy and dissipation"""
    N = 1024
    kw = dict(nu=1e-3, dt=1e-4, force_eps=0.1)
    res = []
    for prec in ("fp64", "fp32bulk"):
        with nsg.Solver(N, precision=prec, **kw) as s:
            s.init("k4", k0=4.0, energy=0.5, seed=11)
            s.step(1)
            res.append(s.measures())
    (ta, ea, _), (tb, eb, _) = res
    sel = ea >= 1e-10 * ea.max()
    assert np.max(np.abs(eb[sel] - ea[sel]) / ea[sel]) < 1e-5
    assert abs(tb.energy - ta.energy) < 1e-7 * ta.energy
    assert abs(tb.dissipation - ta.dissipation) < 1e-6 * ta.dissipation


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32bulk", 1e-6)])
def test_rk5cn_beltrami_decay_512(nsg, precision, tol):
    """the 512^3 Dormand-Prince-family pass 5 on the Beltrami field:
    NL = 0, so every RK5-CN step multiplies the state by (2 - D)/(2 + D)"""
    N, nu, dt = 512, 1e-3, 1e-3
    with nsg.Solver(N, nu=nu, dt=dt, scheme="rk5cn", precision=precision) as s:
        s.init("shapiro")
        u0 = s.get_state()
        s.step(2)
        got = s.get_state()
    D = nu * 12 * dt
    assert rel(got, u0 * ((2 - D) / (2 + D)) ** 2) < tol


def test_dormand_prince_pass5_512(nsg):
    """the 512^3 Dormand-Prince-family pass 5 (FP32 in paired 64-byte row segments, FP64 in whole
    ones) from the same k^4 state: one RK5-CN step in FP32-bulk agrees with FP64 in E(k) to 1e-5 on the shells
    above the FP32 noise floor; one DP5 step (FP64, accepted at once) lands on the RK5-CN state
    (same 5th-order weights), and its error norm, reduced from the per-block sums, sets the next dt"""
    N = 512
    kw = dict(nu=1e-3, dt=1e-3)
    res = {}
    for scheme, prec in (("rk5cn", "fp64"), ("rk5cn", "fp32bulk"), ("dp5", "fp64")):
        with nsg.Solver(N, precision=prec, scheme=scheme, **kw) as s:
            s.init("k4", k0=2.0, energy=0.5, seed=2)
            s.step(1)
            res[scheme, prec] = (s.measures(), s.time(), s.dp_stats() if scheme == "dp5" else None)
    (ta, ea, _), _, _ = res["rk5cn", "fp64"]
    (tb, eb, _), _, _ = res["rk5cn", "fp32bulk"]
    (tc, ec, _), (t, dtn, _), (acc, rej, err) = res["dp5", "fp64"]
    sel = ea >= 1e-10 * ea.max()
    assert np.max(np.abs(eb[sel] - ea[sel]) / ea[sel]) < 1e-5
    assert abs(tb.energy - ta.energy) < 1e-7 * ta.energy
    assert acc == 1 and rej == 0 and t == 1e-3
    assert abs(tc.energy - ta.energy) < 1e-12 * ta.energy
    assert np.max(np.abs(ec[sel] - ea[sel]) / ea[sel]) < 1e-10
    assert 0.0 < err <= 1.0
    assert abs(dtn - 1e-3 * min(1.5, max(0.01, 0.8 * (1.0 / err) ** 0.2))) < 1e-14


def test_determinism_512(nsg):
    """two runs from the same state are bitwise identical (no order-dependent reductions)"""
    N = 512
    out = []
    for _ in range(2):
        with nsg.Solver(N, nu=1e-3, dt=1e-3, cfl=0.5, force_eps=0.1) as s:
            s.init("k4", k0=2.0, energy=0.5, seed=3)
            s.step(2)
            tot, ek, _ = s.measures()
            out.append((s.get_state(), s.time(), np.array(tot), ek))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2]) and np.array_equal(out[0][3], out[1][3])


def test_capacity_1024_fp64(nsg):
    """the 1024^3 FP64 capacity run (about 98 GB of device memory): the corrected Shapiro field is
    Beltrami (nonlinear term zero), so its energy decays by exactly ((2-D)/(2+D))^(2n) under RK4-CN
    and E, Ens, Hel keep the ratios of one shell |k|^2 = 12; a k^4 field with nu = 0 keeps its
    energy (sum Re(u* . NL) = 0) to the RK4 truncation error"""
    N, nu, dt = 1024, 1e-3, 1e-3
    with nsg.Solver(N, nu=nu, dt=dt) as s:
        s.init("shapiro")
        t0 = s.measures()[0]
        s.step(2)
        t1 = s.measures()[0]
        D = nu * 12 * dt
        assert abs(t1.energy / t0.energy - ((2 - D) / (2 + D)) ** 4) < 1e-12
        assert abs(t1.enstrophy / t1.energy - 12.0) < 1e-10
        assert abs(t1.helicity / t1.energy - 12 ** 0.5) < 1e-10
    with nsg.Solver(N, nu=0.0, dt=1e-4) as s:
        s.init("k4", k0=4.0, energy=0.5, seed=11)
        e0 = s.measures()[0].energy
        s.step(2)
        e1 = s.measures()[0].energy
        assert np.isfinite(e1) and abs(e1 - e0) < 1e-9 * e0
