"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so) -- test infrastructure only.

Arrays follow the reference layouts: spectral complex128 (N, N, N//2+1, 3); real float64
(N, N, N, 3); single-component transforms on (N, N, N) <-> (N, N, N//2+1).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    lib = ctypes.CDLL(ORACLE_SO)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.oracle_n_spect.argtypes = [i]
    lib.oracle_n_spect.restype = i
    lib.oracle_set_threads.argtypes = [i]
    lib.oracle_r2c3d.argtypes = [i, vp, vp]
    lib.oracle_c2r3d.argtypes = [i, vp, vp]
    lib.oracle_fft1d.argtypes = [i, vp, i]
    lib.oracle_dealias.argtypes = [i, vp]
    lib.oracle_curl.argtypes = [i, vp, vp]
    lib.oracle_rhs.argtypes = [i, vp, vp]
    lib.oracle_forcing.argtypes = [i, vp, d, d, vp]
    lib.oracle_step.argtypes = [i, vp, d, i, d, i, d, d]
    lib.oracle_measures.argtypes = [i, vp, d, i, i, vp, vp, vp]
    lib.oracle_realspace.argtypes = [i, vp, i, vp]
    lib.oracle_max_speed_sum.argtypes = [i, vp]
    lib.oracle_max_speed_sum.restype = d
    lib.oracle_cfl_dt.argtypes = [i, vp, d]
    lib.oracle_cfl_dt.restype = d
    lib.oracle_ic_taylor_green.argtypes = [i, vp]
    lib.oracle_ic_shapiro.argtypes = [i, d, d, vp]
    lib.oracle_ic_random_ref.argtypes = [i, vp]
    lib.oracle_ic_k4.argtypes = [i, d, d, ctypes.c_uint64, vp]
    lib.oracle_philox4x32_10.argtypes = [vp, vp, vp]
    lib.oracle_step_rk5.argtypes = [i, vp, vp, d, i, d, i, i, d, d]
    lib.oracle_step_rk5.restype = d
    lib.oracle_dp5_factor.argtypes = [d]
    lib.oracle_dp5_factor.restype = d
    lib.oracle_dp5_solve.argtypes = [i, vp, d, i, vp, vp, i, d, d, vp, vp]
    lib.oracle_dp5_solve.restype = i
    lib.oracle_incr_list.argtypes = [i, vp]
    lib.oracle_incr_moments.argtypes = [i, vp, vp]
    lib.oracle_incr_ranges.argtypes = [vp, vp]
    lib.oracle_incr_histogram.argtypes = [i, vp, vp, vp]
    lib.oracle_incr_histogram.restype = ctypes.c_long
    lib.oracle_incr_stats.argtypes = [vp, vp]
    _lib = lib
    return lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def spec_shape(N):
    return (N, N, N // 2 + 1, 3)


def r2c3d(x):
    N = x.shape[0]
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty((N, N, N // 2 + 1), dtype=np.complex128)
    load().oracle_r2c3d(N, _p(x), _p(out))
    return out


def c2r3d(X):
    N = X.shape[0]
    X = np.ascontiguousarray(X, dtype=np.complex128)
    out = np.empty((N, N, N), dtype=np.float64)
    load().oracle_c2r3d(N, _p(X), _p(out))
    return out


def fft1d(x, sign):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    load().oracle_fft1d(a.shape[0], _p(a), int(sign))
    return a


def dealias(uh):
    a = np.ascontiguousarray(uh, dtype=np.complex128).copy()
    load().oracle_dealias(a.shape[0], _p(a))
    return a


def curl(uh):
    a = np.ascontiguousarray(uh, dtype=np.complex128)
    out = np.empty_like(a)
    load().oracle_curl(a.shape[0], _p(a), _p(out))
    return out


def rhs(uh, force_eps=0.0, force_kf=2.5):
    a = np.ascontiguousarray(uh, dtype=np.complex128)
    out = np.empty_like(a)
    load().oracle_rhs(a.shape[0], _p(a), _p(out))
    if force_eps:
        load().oracle_forcing(a.shape[0], _p(a), float(force_eps), float(force_kf), _p(out))
    return out


def step(uh, nu, dt, visc_pow=1, scheme="rk4cn", force_eps=0.0, force_kf=2.5, nsteps=1):
    a = np.ascontiguousarray(uh, dtype=np.complex128).copy()
    sc = {"rk4cn": 0, "ifrk4": 1, "rk5cn": 2, "dp5": 3}[scheme]
    for _ in range(nsteps):
        load().oracle_step(a.shape[0], _p(a), float(nu), int(visc_pow), float(dt), sc, float(force_eps),
                           float(force_kf))
    return a


def step_rk5(uh, nu, dt, nstage=6, visc_pow=1, quirk=False, force_eps=0.0, force_kf=2.5):
    """one Dormand-Prince RK step; returns (u5, embedded error norm or 0 for nstage 6)"""
    a = np.ascontiguousarray(uh, dtype=np.complex128)
    out = np.empty_like(a)
    err = load().oracle_step_rk5(a.shape[0], _p(a), _p(out), float(nu), int(visc_pow), float(dt), int(nstage),
                                 int(quirk), float(force_eps), float(force_kf))
    return out, err


def dp5_factor(err):
    return load().oracle_dp5_factor(float(err))


def dp5_solve(uh, nu, dt, nsteps, t=0.0, visc_pow=1, force_eps=0.0, force_kf=2.5):
    """nsteps accepted adaptive Dormand-Prince steps: (u, t, dt_next, rejected, last_err, rc)"""
    a = np.ascontiguousarray(uh, dtype=np.complex128).copy()
    dtv, tv = np.array([dt], dtype=np.float64), np.array([t], dtype=np.float64)
    rej, last = np.zeros(1, dtype=np.int64), np.zeros(1)
    rc = load().oracle_dp5_solve(a.shape[0], _p(a), float(nu), int(visc_pow), _p(dtv), _p(tv), int(nsteps),
                                 float(force_eps), float(force_kf), _p(rej), _p(last))
    return a, float(tv[0]), float(dtv[0]), int(rej[0]), float(last[0]), rc


def measures(uh, nu, visc_pow=1, quirk=False):
    a = np.ascontiguousarray(uh, dtype=np.complex128)
    N = a.shape[0]
    ns = load().oracle_n_spect(N)
    tot = np.zeros(5)
    ek = np.zeros(ns)
    ens = np.zeros(ns)
    load().oracle_measures(N, _p(a), float(nu), int(visc_pow), int(quirk), _p(tot), _p(ek), _p(ens))
    return tot, ek, ens


def realspace(uh, which=0):
    a = np.ascontiguousarray(uh, dtype=np.complex128)
    N = a.shape[0]
    out = np.empty((N, N, N, 3))
    load().oracle_realspace(N, _p(a), int(which), _p(out))
    return out


def max_speed_sum(uh):
    a = np.ascontiguousarray(uh, dtype=np.complex128)
    return load().oracle_max_speed_sum(a.shape[0], _p(a))


def cfl_dt(uh, cfl):
    a = np.ascontiguousarray(uh, dtype=np.complex128)
    return load().oracle_cfl_dt(a.shape[0], _p(a), float(cfl))


def ic_taylor_green(N):
    out = np.empty(spec_shape(N), dtype=np.complex128)
    load().oracle_ic_taylor_green(N, _p(out))
    return out


def ic_shapiro(N, nu=0.0, t=0.0):
    out = np.empty(spec_shape(N), dtype=np.complex128)
    load().oracle_ic_shapiro(N, float(nu), float(t), _p(out))
    return out


def ic_random_ref(N):
    out = np.empty(spec_shape(N), dtype=np.complex128)
    load().oracle_ic_random_ref(N, _p(out))
    return out


def ic_k4(N, k0=4.0, energy=0.5, seed=1):
    out = np.empty(spec_shape(N), dtype=np.complex128)
    load().oracle_ic_k4(N, float(k0), float(energy), int(seed), _p(out))
    return out


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    load().oracle_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def n_spect(N):
    return load().oracle_n_spect(N)


def set_threads(n):
    load().oracle_set_threads(int(n))


# ---------------------------------------------------------------- increment statistics
NUM_INCR, N_BINS = 9, 1000


def incr_list(N):
    r = np.zeros(NUM_INCR, dtype=np.int32)
    load().oracle_incr_list(int(N), _p(r))
    return r


def incr_moments(u, acc=None):
    """one Precompute pass over a real field u (N, N, N, 3); acc [9, 7] = n, min, max, mean, M2, M3, M4"""
    u = np.ascontiguousarray(u, dtype=np.float64)
    acc = np.zeros((NUM_INCR, 7)) if acc is None else acc
    load().oracle_incr_moments(u.shape[0], _p(u), _p(acc))
    return acc


def incr_ranges(acc):
    acc = np.ascontiguousarray(acc, dtype=np.float64)
    out = np.empty((NUM_INCR, N_BINS + 1))
    load().oracle_incr_ranges(_p(acc), _p(out))
    return out


def incr_histogram(u, ranges, counts=None):
    """one ComputeStats pass: (counts [9, 1000] accumulated, number of values outside the ranges)"""
    u = np.ascontiguousarray(u, dtype=np.float64)
    ranges = np.ascontiguousarray(ranges, dtype=np.float64)
    counts = np.zeros((NUM_INCR, N_BINS)) if counts is None else counts
    out = load().oracle_incr_histogram(u.shape[0], _p(u), _p(ranges), _p(counts))
    return counts, int(out)


def incr_stats(acc):
    acc = np.ascontiguousarray(acc, dtype=np.float64)
    out = np.empty((NUM_INCR, 6))
    load().oracle_incr_stats(_p(acc), _p(out))
    return out
