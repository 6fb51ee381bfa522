#!/usr/bin/env python3
"""Writes tests/golden/*.npz: golden vectors for the CPU oracle and the GPU library.

TEST INFRASTRUCTURE.  This is an independent numpy restatement of the reference time-step path
(numpy.fft = pocketfft for every transform, vectorised formulas instead of the oracle's loops):

  * wavenumbers          Solver/src/solver.c:1764-1826 (kx, ky Nyquist stored as +N/2)
  * dealias (2/3 rule)   Solver/src/solver.c:1709-1756 (|k| > floor(N/3) zeroed; D2 corrected)
  * NonlinearRHSBatch    Solver/src/solver.c:620-731   (D1 corrected: transforms of size N^3)
  * RK4Step (RK4-CN)     Solver/src/solver.c:505-608, coefficients :28-32
  * measures + spectra   Solver/src/solver.c:1142-1276 (correct form and the D3/D14 quirk form)
  * IF-RK4, constant-power forcing and the Philox k^4 initial condition (no reference counterpart)

Run:  python tests/golden/make_golden.py   (seconds; output is small and committed)
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


# ----------------------------------------------------------------------------------------------
# grid
# ----------------------------------------------------------------------------------------------
def wavenumbers(N):
    k = np.array([i if i <= N // 2 else i - N for i in range(N)], dtype=np.int64)
    kz = np.arange(N // 2 + 1, dtype=np.int64)
    KX, KY, KZ = np.meshgrid(k, k, kz, indexing="ij")
    return KX, KY, KZ


def dealias(uh):
    N = uh.shape[0]
    KX, KY, KZ = wavenumbers(N)
    mask = np.sqrt((KX * KX + KY * KY + KZ * KZ).astype(float)) > (N // 3)
    out = uh.copy()
    out[mask] = 0.0
    return out


def r2c(u):  # FFTW-normalised forward transform of a (N,N,N,3) real field -> (N,N,Nzf,3)
    return np.fft.rfftn(u, axes=(0, 1, 2))


def c2r(uh):  # unnormalised backward transform (FFTW c2r convention)
    N = uh.shape[0]
    return np.fft.irfftn(uh, s=(N, N, N), axes=(0, 1, 2)) * N ** 3


def curl(uh):
    KX, KY, KZ = wavenumbers(uh.shape[0])
    ux, uy, uz = uh[..., 0], uh[..., 1], uh[..., 2]
    return 1j * np.stack([KY * uz - KZ * uy, KZ * ux - KX * uz, KX * uy - KY * ux], axis=-1)


def project(nl):
    KX, KY, KZ = wavenumbers(nl.shape[0])
    k2 = (KX * KX + KY * KY + KZ * KZ).astype(float)
    kd = KX * nl[..., 0] + KY * nl[..., 1] + KZ * nl[..., 2]
    inv = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1.0), 0.0)
    out = nl - np.stack([KX, KY, KZ], axis=-1) * (kd * inv)[..., None]
    out[k2 == 0] = 0.0
    return out


def rhs(uh):
    N = uh.shape[0]
    u = c2r(uh)
    w = c2r(curl(uh))
    c = np.cross(u, w)
    nl = r2c(c) / float(N ** 3) ** 2
    return dealias(project(nl))


def forcing(uh, eps_in, kf):
    N = uh.shape[0]
    KX, KY, KZ = wavenumbers(N)
    k2 = KX * KX + KY * KY + KZ * KZ
    band = (k2 > 0) & (k2 <= kf * kf)
    wgt = np.where((KZ == 0) | (KZ == N // 2), 1.0, 2.0)
    Ef = 0.5 * np.sum(wgt[band][:, None] * np.abs(uh[band]) ** 2) / float(N) ** 6
    f = np.zeros_like(uh)
    if Ef > 0:
        f[band] = eps_in / (2.0 * Ef) * uh[band]
    return f


def visc(N, nu, p):
    KX, KY, KZ = wavenumbers(N)
    return nu * (KX * KX + KY * KY + KZ * KZ).astype(float) ** p


def step(uh, nu, dt, p=1, scheme="rk4cn", eps=0.0, kf=2.5):
    N = uh.shape[0]
    F = (lambda v: rhs(v) + forcing(v, eps, kf)) if eps else rhs
    D = visc(N, nu, p)[..., None]
    k1 = F(uh)
    if scheme == "rk4cn":
        k2 = F(uh + 0.5 * dt * k1)
        k3 = F(uh + 0.5 * dt * k2)
        k4 = F(uh + dt * k3)
        Dd = D * dt
        return uh * ((2 - Dd) / (2 + Dd)) + (2 * dt / (2 + Dd)) * (k1 / 6 + k2 / 3 + k3 / 3 + k4 / 6)
    E, E2 = np.exp(-D * dt), np.exp(-D * dt / 2)
    k2 = F(E2 * (uh + 0.5 * dt * k1))
    k3 = F(E2 * uh + 0.5 * dt * k2)
    k4 = F(E * uh + dt * E2 * k3)
    return E * uh + dt / 6 * (E * k1 + 2 * E2 * (k2 + k3) + k4)


def measures(uh, nu, p=1, quirk=False):
    N = uh.shape[0]
    KX, KY, KZ = wavenumbers(N)
    wh = curl(uh)
    ch = curl(wh)
    au, aw, ac = np.abs(uh) ** 2, np.abs(wh) ** 2, np.abs(ch) ** 2
    edge = (KZ == 0) | (KZ == N // 2)
    pre = visc(N, nu, p)
    if quirk:
        hel = np.real(np.sum(uh * wh, axis=-1))
        dbl = lambda a: np.where(edge, a.sum(-1), 2 * a[..., 0] + a[..., 1] + a[..., 2])
    else:
        hel = np.real(np.sum(np.conj(uh) * wh, axis=-1))
        dbl = lambda a: np.where(edge, 1.0, 2.0) * a.sum(-1)
    w2 = np.where(edge, 1.0, 2.0)
    nf, cf = 0.5 / float(N) ** 6, 8 * np.pi ** 3
    tot = np.array([dbl(au).sum(), dbl(aw).sum(), dbl(ac).sum(), (w2 * hel).sum(),
                    (w2 * pre * au.sum(-1)).sum()])
    tot *= cf * nf
    tot[4] *= 2.0
    ns = int(np.sqrt(3 * (N / 2.0) ** 2)) + 1
    b = np.rint(np.sqrt((KX * KX + KY * KY + KZ * KZ).astype(float))).astype(int)
    ek = np.zeros(ns)
    ens = np.zeros(ns)
    ok = b < ns
    np.add.at(ek, b[ok], (cf * nf * w2 * au.sum(-1))[ok])
    np.add.at(ens, b[ok], (cf * nf * w2 * aw.sum(-1))[ok])
    return tot, ek, ens


# ----------------------------------------------------------------------------------------------
# Philox-4x32-10 and the k^4 random-phase initial condition
# ----------------------------------------------------------------------------------------------
M32 = np.uint64(0xFFFFFFFF)


def philox(c0, c1, c2, c3, k0, k1):
    c = [np.asarray(x, dtype=np.uint64) & M32 for x in (c0, c1, c2, c3)]
    k0 = np.uint64(k0) & M32
    k1 = np.uint64(k1) & M32
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c[0]
        p1 = np.uint64(0xCD9E8D57) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & M32
        hi1, lo1 = p1 >> np.uint64(32), p1 & M32
        c = [(hi1 ^ c[1] ^ k0) & M32, lo1, (hi0 ^ c[3] ^ k1) & M32, lo0]
        k0 = (k0 + np.uint64(0x9E3779B9)) & M32
        k1 = (k1 + np.uint64(0xBB67AE85)) & M32
    return c


def u01(a, b):
    v = ((a << np.uint64(21)) ^ (b >> np.uint64(11))) & np.uint64((1 << 53) - 1)
    return (v.astype(np.float64) + 0.5) / 9007199254740992.0


def k4_ic(N, k0, energy, seed):
    KX, KY, KZ = wavenumbers(N)
    # mirror the kz=0 half plane: (kx,ky,0) with ky<0 or (ky==0 and kx<0) takes conj of (-kx,-ky,0)
    mir = (KZ == 0) & ((KY < 0) | ((KY == 0) & (KX < 0)))
    gx, gy = np.where(mir, -KX, KX), np.where(mir, -KY, KY)
    k2 = gx * gx + gy * gy + KZ * KZ
    i = np.where(gx >= 0, gx, gx + N).astype(np.uint64)
    j = np.where(gy >= 0, gy, gy + N).astype(np.uint64)
    n = (i * np.uint64(N) + j) * np.uint64(N // 2 + 1) + KZ.astype(np.uint64)
    lo, hi = n & M32, n >> np.uint64(32)
    r0 = philox(lo, hi, 0, 0, seed & 0xFFFFFFFF, seed >> 32)
    r1 = philox(lo, hi, 1, 0, seed & 0xFFFFFFFF, seed >> 32)
    th1, th2, phi = 2 * np.pi * u01(r0[0], r0[1]), 2 * np.pi * u01(r0[2], r0[3]), 2 * np.pi * u01(r1[0], r1[1])
    kk = np.sqrt(k2.astype(float))
    keep = (k2 > 0) & (k2 <= (N // 3) ** 2)
    kk_s = np.where(keep, kk, 1.0)
    Ek = k2.astype(float) ** 2 * np.exp(-2 * (kk_s / k0) ** 2)
    A = np.sqrt(Ek / (4 * np.pi * np.where(keep, k2, 1)))
    al = A * np.exp(1j * th1) * np.cos(phi)
    be = A * np.exp(1j * th2) * np.sin(phi)
    kh2 = gx * gx + gy * gy
    kh = np.sqrt(np.where(kh2 > 0, kh2, 1).astype(float))
    e1 = np.where(kh2[..., None] > 0, np.stack([gy / kh, -gx / kh, 0 * kh], -1), np.array([1.0, 0, 0]))
    e2 = np.where(kh2[..., None] > 0, np.stack([gx * KZ / (kk_s * kh), gy * KZ / (kk_s * kh), -kh2 / (kk_s * kh)], -1),
                  np.array([0.0, 1.0, 0]))
    uh = al[..., None] * e1 + be[..., None] * e2
    uh[~keep] = 0.0
    uh = np.where(mir[..., None], np.conj(uh), uh)
    wgt = np.where((KZ == 0) | (KZ == N // 2), 1.0, 2.0)
    e_now = 0.5 * np.sum(wgt[..., None] * np.abs(uh) ** 2)
    return uh * (np.sqrt(energy / e_now) * N ** 3)


# ----------------------------------------------------------------------------------------------
def random_state(N, seed):
    """dealiased, Hermitian state from a seeded real field"""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((N, N, N, 3))
    return dealias(r2c(u))


def main():
    out = {}
    for N in (8, 12, 16, 24):
        x = np.random.default_rng(100 + N).standard_normal((N, N, N))
        out[f"fft_in_{N}"] = x
        out[f"fft_out_{N}"] = np.fft.rfftn(x)
    np.savez_compressed(os.path.join(HERE, "fft.npz"), **out)

    out = {}
    for N in (12, 16):
        uh = random_state(N, 200 + N)
        out[f"rhs_in_{N}"] = uh
        out[f"rhs_out_{N}"] = rhs(uh)
        out[f"rhs_forced_out_{N}"] = rhs(uh) + forcing(uh, 0.1, 2.5)
    N = 16
    uh = random_state(N, 300) * 0.25
    out["step_in_16"] = uh
    out["step_rk4cn_16"] = step(uh, 0.05, 1e-2, 1, "rk4cn")
    out["step_rk4cn_hyper_16"] = step(uh, 1e-3, 1e-2, 2, "rk4cn")
    out["step_ifrk4_16"] = step(uh, 0.05, 1e-2, 1, "ifrk4")
    out["step_ifrk4_forced_16"] = step(uh, 0.05, 1e-2, 1, "ifrk4", eps=0.1, kf=2.5)
    tot, ek, ens = measures(uh, 0.05, 1, False)
    out["meas_tot_16"], out["meas_ek_16"], out["meas_ens_16"] = tot, ek, ens
    tot, ek, ens = measures(uh, 0.05, 1, True)
    out["meas_quirk_tot_16"] = tot
    out["k4_16"] = k4_ic(16, 2.0, 0.5, 7)
    out["k4_24"] = k4_ic(24, 3.0, 0.5, (1 << 40) + 5)
    np.savez_compressed(os.path.join(HERE, "physics.npz"), **out)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
