"""Oracle physics: golden vectors (numpy restatement), closed forms and invariants.

Reference lines: NonlinearRHSBatch Solver/src/solver.c:620-731, RK4Step :505-608, dealias
:1709-1756, measures :1142-1276, ICs :1537-1648.
"""
import numpy as np
import pytest

from conftest import rel, wavenumbers

NU_TG = 1.0 / 1600.0


# ---------------------------------------------------------------------------- golden vectors
@pytest.mark.parametrize("N", [12, 16])
def test_rhs_golden(oracle, golden, N):
    g = golden["phys"]
    assert rel(oracle.rhs(g[f"rhs_in_{N}"]), g[f"rhs_out_{N}"]) < 1e-13
    assert rel(oracle.rhs(g[f"rhs_in_{N}"], 0.1, 2.5), g[f"rhs_forced_out_{N}"]) < 1e-13


@pytest.mark.parametrize("name,args", [("rk4cn", (0.05, 1e-2, 1, "rk4cn")), ("rk4cn_hyper", (1e-3, 1e-2, 2, "rk4cn")),
                                       ("ifrk4", (0.05, 1e-2, 1, "ifrk4"))])
def test_step_golden(oracle, golden, name, args):
    g = golden["phys"]
    u0 = g["step_in_16"]
    out = oracle.step(u0, *args)
    # compare the increment, which carries the nonlinear term
    assert rel(out - u0, g[f"step_{name}_16"] - u0) < 1e-13


def test_step_forced_golden(oracle, golden):
    g = golden["phys"]
    u0 = g["step_in_16"]
    out = oracle.step(u0, 0.05, 1e-2, 1, "ifrk4", 0.1, 2.5)
    assert rel(out - u0, g["step_ifrk4_forced_16"] - u0) < 1e-13


def test_measures_golden(oracle, golden):
    g = golden["phys"]
    tot, ek, ens = oracle.measures(g["step_in_16"], 0.05, 1, False)
    assert np.allclose(tot, g["meas_tot_16"], rtol=1e-13, atol=0)
    assert rel(ek, g["meas_ek_16"]) < 1e-13 and rel(ens, g["meas_ens_16"]) < 1e-13
    totq, _, _ = oracle.measures(g["step_in_16"], 0.05, 1, True)
    gq = g["meas_quirk_tot_16"]
    assert np.allclose(totq[[0, 1, 2, 4]], gq[[0, 1, 2, 4]], rtol=1e-13, atol=0)
    # conjugate-free helicity vanishes for a Hermitian field: both values are round-off
    assert abs(totq[3]) < 1e-12 * abs(tot[1]) and abs(gq[3]) < 1e-12 * abs(tot[1])


@pytest.mark.parametrize("N,k0,seed", [(16, 2.0, 7), (24, 3.0, (1 << 40) + 5)])
def test_k4_ic_golden(oracle, golden, N, k0, seed):
    assert rel(oracle.ic_k4(N, k0, 0.5, seed), golden["phys"][f"k4_{N}"]) < 1e-13


def test_philox_known_answers(oracle):
    # Random123 known-answer vectors for philox4x32-10
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in kat:
        assert tuple(int(v) for v in oracle.philox(ctr, key)) == want


# ---------------------------------------------------------------------------- closed forms
def test_taylor_green_measures(oracle):
    """TG at t=0 (Solver/src/solver.c:1553-1574): E=pi^3, Ens=3pi^3, Pal=9pi^3, Hel=0, eps=6 nu pi^3;
    E(k) only in bin round(sqrt 3)=2; reference quirk arithmetic gives 0.75, 1.75, 6.75 pi^3."""
    N = 32
    uh = oracle.ic_taylor_green(N)
    tot, ek, ens = oracle.measures(uh, NU_TG, 1, False)
    p3 = np.pi ** 3
    assert np.allclose(tot[[0, 1, 2, 4]], [p3, 3 * p3, 9 * p3, 6 * NU_TG * p3], rtol=1e-13)
    assert abs(tot[3]) < 1e-12
    assert abs(ek[2] - p3) < 1e-12 * p3 and abs(np.sum(ek) - ek[2]) < 1e-12
    assert abs(ens[2] - 3 * p3) < 1e-12 * p3
    totq, _, _ = oracle.measures(uh, NU_TG, 1, True)
    assert np.allclose(totq[:3], [0.75 * p3, 1.75 * p3, 6.75 * p3], rtol=1e-13)


def test_taylor_green_rhs_closed_form(oracle):
    """NL of TG built from the analytic u and w = curl u on the grid, then r2c/N^6, projection and
    dealias (numpy), against the oracle's NonlinearRHSBatch of the TG initial condition."""
    N = 16
    x = 2 * np.pi * np.arange(N) / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), 0 * X], -1)
    w = np.stack([-np.cos(X) * np.sin(Y) * np.sin(Z), -np.sin(X) * np.cos(Y) * np.sin(Z),
                  2 * np.sin(X) * np.sin(Y) * np.cos(Z)], -1)
    # the reference forms c2r(u_hat) c2r(w_hat) / N^6 = u x w of the physical fields
    nl = np.fft.rfftn(np.cross(u, w), axes=(0, 1, 2))
    KX, KY, KZ = wavenumbers(N)
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    kd = KX * nl[..., 0] + KY * nl[..., 1] + KZ * nl[..., 2]
    nl = nl - np.stack([KX, KY, KZ], -1) * np.where(k2 > 0, kd / np.where(k2 > 0, k2, 1), 0)[..., None]
    nl[(k2 == 0) | (np.sqrt(k2) > N // 3)] = 0
    got = oracle.rhs(oracle.ic_taylor_green(N))
    assert np.max(np.abs(got - nl)) < 1e-12 * np.max(np.abs(nl))


def test_shapiro_is_beltrami_and_nl_vanishes(oracle):
    """Shapiro with the divergence-free sin(Mz) (D17 corrected) satisfies w = +lambda u, so
    u x w = 0 and the nonlinear term vanishes to round-off."""
    N = 16
    uh = oracle.ic_shapiro(N)
    lam = np.sqrt(12.0)
    assert np.max(np.abs(oracle.curl(uh) - lam * uh)) < 1e-13 * np.max(np.abs(uh)) * lam
    KX, KY, KZ = wavenumbers(N)
    assert np.max(np.abs(KX * uh[..., 0] + KY * uh[..., 1] + KZ * uh[..., 2])) < 1e-10
    assert np.max(np.abs(oracle.rhs(uh))) < 1e-13 * np.max(np.abs(uh)) ** 2 * lam / N ** 3
    tot, _, _ = oracle.measures(uh, 0.01, 1, False)
    assert abs(tot[1] / tot[0] - lam ** 2) < 1e-12 * lam ** 2  # Ens/E
    assert abs(tot[2] / tot[0] - lam ** 4) < 1e-12 * lam ** 4  # Pal/E
    assert abs(tot[3] / tot[0] - lam) < 1e-12 * lam  # Hel/E


@pytest.mark.parametrize("scheme", ["rk4cn", "ifrk4"])
def test_shapiro_exact_decay(oracle, scheme):
    """NL = 0 for the Beltrami field, so RK4-CN multiplies every mode by (2-D)/(2+D) per step,
    D = nu lambda^2 dt, and IF-RK4 by exp(-nu lambda^2 dt)."""
    N, nu, dt, n = 12, 0.01, 0.01, 5
    u0 = oracle.ic_shapiro(N)
    got = oracle.step(u0, nu, dt, 1, scheme, nsteps=n)
    D = nu * 12.0 * dt
    fac = ((2 - D) / (2 + D)) ** n if scheme == "rk4cn" else np.exp(-12.0 * nu * dt * n)
    assert rel(got, u0 * fac) < 1e-13


def test_ifrk4_fourth_order(oracle):
    """Self-convergence of IF-RK4 on TG (N=12): error ratio for dt -> dt/2 is about 2^4."""
    N, nu, T = 12, 0.05, 0.4
    u0 = oracle.ic_taylor_green(N) * 3.0
    ref = oracle.step(u0, nu, T / 64, 1, "ifrk4", nsteps=64)
    e1 = rel(oracle.step(u0, nu, T / 4, 1, "ifrk4", nsteps=4), ref)
    e2 = rel(oracle.step(u0, nu, T / 8, 1, "ifrk4", nsteps=8), ref)
    assert 12.0 < e1 / e2 < 20.0


def _rk4_planted(oracle, u, dt, n, a43):
    """numpy RK4 over the oracle's nonlinear term with the last stage's input coefficient a43
    (1 in RK4; a planted mistake otherwise), nu = 0"""
    for _ in range(n):
        k1 = oracle.rhs(u)
        k2 = oracle.rhs(u + 0.5 * dt * k1)
        k3 = oracle.rhs(u + 0.5 * dt * k2)
        k4 = oracle.rhs(u + a43 * dt * k3)
        u = u + dt * (k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6)
    return u


def test_rk4cn_fourth_order_inviscid(oracle):
    """RK4Step (Solver/src/solver.c:505-608) with nu = 0: D = 0 makes the Crank-Nicolson factor 1
    and the scheme is the classical RK4 on the nonlinear term, so the self-convergence error ratio
    for dt -> dt/2 is about 2^4.  The stage coefficients (1/2, 1/2, 1) and weights (1/6, 1/3, 1/3,
    1/6) are all exercised: the same test on a copy with the last stage's input at 0.5 dt instead
    of dt (a plausible slip) drops to second order, and the oracle equals the classical scheme run
    over its own nonlinear term."""
    N, T = 12, 0.4
    u0 = oracle.ic_taylor_green(N) * 3.0
    u0 = u0 + 0.3 * oracle.ic_k4(N, 2.0, 0.5, 21)  # break the TG symmetries
    ref = oracle.step(u0, 0.0, T / 64, 1, "rk4cn", nsteps=64)
    e1 = rel(oracle.step(u0, 0.0, T / 4, 1, "rk4cn", nsteps=4), ref)
    e2 = rel(oracle.step(u0, 0.0, T / 8, 1, "rk4cn", nsteps=8), ref)
    e3 = rel(oracle.step(u0, 0.0, T / 16, 1, "rk4cn", nsteps=16), ref)
    assert 13.0 < e1 / e2 < 19.0 and 13.0 < e2 / e3 < 19.0
    assert rel(oracle.step(u0, 0.0, T / 8, 1, "rk4cn", nsteps=8), _rk4_planted(oracle, u0, T / 8, 8, 1.0)) < 1e-13
    ref_w = _rk4_planted(oracle, u0, T / 64, 64, 0.5)
    w1 = rel(_rk4_planted(oracle, u0, T / 4, 4, 0.5), ref_w)
    w2 = rel(_rk4_planted(oracle, u0, T / 8, 8, 0.5), ref_w)
    assert w1 / w2 < 9.0


def test_rk4cn_viscous_factor(oracle):
    """single mode of a linear problem: a Beltrami field has NL = 0, so one RK4-CN step multiplies
    the mode by (2 - D)/(2 + D) with D = dt nu |k|^(2p) (Solver/src/solver.c:590-603) -- for both
    viscosity powers, which separates the factor from e^{-D} and from (1 - D/2)/(1 + D/2)^2"""
    N, dt = 12, 0.05
    u0 = oracle.ic_shapiro(N)
    for p, nu in [(1, 0.3), (2, 0.02)]:
        D = dt * nu * 12.0 ** p
        got = oracle.step(u0, nu, dt, p, "rk4cn")
        assert rel(got, u0 * (2 - D) / (2 + D)) < 1e-13


# ---------------------------------------------------------------------------- invariants
@pytest.mark.parametrize("N", [12, 16])
def test_rhs_invariants(oracle, golden, N):
    uh = golden["phys"][f"rhs_in_{N}"]
    nl = oracle.rhs(uh)
    KX, KY, KZ = wavenumbers(N)
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    assert np.max(np.abs(KX * nl[..., 0] + KY * nl[..., 1] + KZ * nl[..., 2])) < 1e-12 * np.max(np.abs(nl)) * N
    assert np.all(nl[np.sqrt(k2) > N // 3] == 0)
    # Hermitian kz=0 plane: NL(-kx,-ky,0) = conj NL(kx,ky,0)
    p = nl[:, :, 0]
    m = np.conj(p[(-np.arange(N)) % N][:, (-np.arange(N)) % N])
    assert np.max(np.abs(p - m)) < 1e-12 * np.max(np.abs(nl))
    # energy conservation of the nonlinear term for a solenoidal field: sum' Re(u* . NL) = 0
    uh = oracle.ic_k4(N, 2.0, 0.5, 11)
    nl = oracle.rhs(uh)
    w = np.where((KZ == 0) | (KZ == N // 2), 1.0, 2.0)
    tr = np.sum(w[..., None] * np.real(np.conj(uh) * nl))
    assert abs(tr) < 1e-11 * np.sum(w[..., None] * np.abs(uh) * np.abs(nl))


@pytest.mark.parametrize("N", [8, 12, 16, 24, 32, 48, 64])
def test_dealias_mode_count(oracle, N):
    """modes kept by the 2/3 rule = lattice points with |k| <= floor(N/3) in the half spectrum"""
    uh = np.ones((N, N, N // 2 + 1, 3), dtype=complex)
    kept = np.count_nonzero(oracle.dealias(uh)[..., 0])
    c = N // 3
    r = np.arange(-c, c + 1)
    brute = sum(1 for a in r for b in r for z in range(0, c + 1) if a * a + b * b + z * z <= c * c)
    assert kept == brute
    assert np.array_equal(oracle.dealias(oracle.dealias(uh)), oracle.dealias(uh))


def test_k4_ic_properties(oracle):
    N = 32
    uh = oracle.ic_k4(N, 4.0, 0.5, 1)
    KX, KY, KZ = wavenumbers(N)
    assert np.max(np.abs(KX * uh[..., 0] + KY * uh[..., 1] + KZ * uh[..., 2])) < 1e-9 * np.max(np.abs(uh))
    p = uh[:, :, 0]
    m = np.conj(p[(-np.arange(N)) % N][:, (-np.arange(N)) % N])
    assert np.max(np.abs(p - m)) == 0.0
    u = oracle.realspace(uh, 0)
    assert abs(0.5 * np.mean(np.sum(u * u, -1)) - 0.5) < 1e-12
    tot, _, _ = oracle.measures(uh, 0.0, 1, False)
    assert abs(tot[0] / (8 * np.pi ** 3) - 0.5) < 1e-12


def test_forcing_injection_rate(oracle, golden):
    """constant power: volume-averaged injection sum' Re(u* . f)/N^6 equals eps_in"""
    N = 16
    uh = golden["phys"]["rhs_in_16"]
    f = oracle.rhs(uh, 0.1, 2.5) - oracle.rhs(uh)
    KX, KY, KZ = wavenumbers(N)
    w = np.where((KZ == 0) | (KZ == N // 2), 1.0, 2.0)
    inj = np.sum(w[..., None] * np.real(np.conj(uh) * f)) / N ** 6
    assert abs(inj - 0.1) < 1e-13
    band = (KX ** 2 + KY ** 2 + KZ ** 2) <= 6.25
    assert np.all(f[~band] == 0)


def test_cfl_dt(oracle, golden):
    uh = golden["phys"]["rhs_in_16"]
    u = oracle.realspace(uh, 0)
    m = np.max(np.sum(np.abs(u), -1))
    assert abs(oracle.max_speed_sum(uh) - m) < 1e-14 * m
    assert abs(oracle.cfl_dt(uh, 0.5) - 0.5 * (2 * np.pi / 16) / m) < 1e-15


def test_realspace_matches_numpy(oracle, golden):
    uh = golden["phys"]["rhs_in_12"]
    u = oracle.realspace(uh, 0)
    assert rel(u, np.fft.irfftn(uh, s=(12, 12, 12), axes=(0, 1, 2))) < 1e-14
    w = oracle.realspace(uh, 1)
    assert rel(w, np.fft.irfftn(oracle.curl(uh), s=(12, 12, 12), axes=(0, 1, 2))) < 1e-14


def test_restart_is_bitwise(oracle, golden):
    """RK4 is one-step: (u_hat, dt) is a complete restart state"""
    u0 = golden["phys"]["step_in_16"]
    a = oracle.step(u0, 0.05, 1e-2, nsteps=3)
    b = oracle.step(oracle.step(u0, 0.05, 1e-2, nsteps=2).copy(), 0.05, 1e-2, nsteps=1)
    assert np.array_equal(a, b)
