"""Python host binding of libnsg.so (the MI355X pseudo-spectral time-step library).

The package directory name starts with a digit, so load it by path::

    import importlib.util, pathlib
    spec = importlib.util.spec_from_file_location(
        "navier_stokes_amd", "3d_navier_stokes_amd/__init__.py",
        submodule_search_locations=["3d_navier_stokes_amd"])

The binding mirrors the reference's run parameters (Solver/src/utils.c:45-67) and its integrator
entry points: ``Solver.step`` is RK4Step (Solver/src/solver.c:505), ``Solver.rhs`` is
NonlinearRHSBatch (:620), ``Solver.measures`` is ComputeSystemMeasurables (:1142).  Spectral arrays
are complex128 of shape (N, N, N//2+1, 3) -- the reference AoS layout; real arrays are float64 of
shape (N, N, N, 3).  Every call goes to the HIP library; there is no CPU fallback: if libnsg.so is
missing or fails to load, ``load_library`` raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import NamedTuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnsg.so")

SCHEMES = {"rk4cn": 0, "ifrk4": 1, "rk5cn": 2, "dp5": 3}
PRECISIONS = {"fp64": 0, "fp32bulk": 1}
ICS = {"taylor_green": 0, "shapiro": 1, "k4": 2}
SUPPORTED_N = (8, 12, 16, 24, 32, 48, 64, 96, 128, 256, 512, 1024)

ERRORS = {-1: "NSG_ERR_INVALID", -2: "NSG_ERR_OOM", -3: "NSG_ERR_DEVICE", -4: "NSG_ERR_NONFINITE",
          -5: "NSG_ERR_BLOWUP", -6: "NSG_ERR_STEP"}


class NSGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class _Params(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int * 3), ("nu", ctypes.c_double), ("visc_pow", ctypes.c_int),
                ("dt", ctypes.c_double), ("t0", ctypes.c_double), ("scheme", ctypes.c_int),
                ("precision", ctypes.c_int), ("cfl", ctypes.c_double), ("force_eps", ctypes.c_double),
                ("force_kf", ctypes.c_double), ("max_speed", ctypes.c_double), ("device", ctypes.c_int)]


class _Totals(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("enstrophy", ctypes.c_double),
                ("palinstrophy", ctypes.c_double), ("helicity", ctypes.c_double),
                ("dissipation", ctypes.c_double)]


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libnsg.so (raises OSError if it is absent: the HIP library is the only backend)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} not found: build it with `make -C 3d_navier_stokes_amd/csrc`")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp, dp = ctypes.c_void_p, P(ctypes.c_double)
    lib.nsg_default_params.argtypes = [P(_Params)]
    lib.nsg_create.argtypes = [P(_Params), P(vp)]
    lib.nsg_destroy.argtypes = [vp]
    lib.nsg_last_error.argtypes = [vp]
    lib.nsg_last_error.restype = ctypes.c_char_p
    lib.nsg_n_spect.argtypes = [vp]
    for name in ("nsg_set_state", "nsg_get_state", "nsg_set_state_real"):
        getattr(lib, name).argtypes = [vp, vp]
    lib.nsg_init_state.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]
    lib.nsg_rhs.argtypes = [vp, vp, vp]
    lib.nsg_step.argtypes = [vp, ctypes.c_long]
    lib.nsg_dp_stats.argtypes = [vp, P(ctypes.c_longlong), P(ctypes.c_longlong), dp]
    lib.nsg_measures.argtypes = [vp, P(_Totals), vp, vp, ctypes.c_int, ctypes.c_int]
    lib.nsg_measures_noalloc.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
    lib.nsg_realspace.argtypes = [vp, ctypes.c_int, vp]
    lib.nsg_get_time.argtypes = [vp, dp, dp, dp]
    lib.nsg_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    lib.nsg_profile_passes.argtypes = [vp, ctypes.c_int, dp, dp]
    lib.nsg_step_bytes.argtypes = [vp]
    lib.nsg_step_bytes.restype = ctypes.c_double
    lib.nsg_schedule_count.argtypes = [ctypes.c_int]
    lib.nsg_set_schedule.argtypes = [vp, ctypes.c_int]
    lib.nsg_autotune.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_int), dp]
    # ensembles
    lib.nsg_ensemble_create.argtypes = [P(_Params), ctypes.c_int, P(vp)]
    lib.nsg_ensemble_destroy.argtypes = [vp]
    lib.nsg_ensemble_last_error.argtypes = [vp]
    lib.nsg_ensemble_last_error.restype = ctypes.c_char_p
    lib.nsg_ensemble_group_count.argtypes = [vp]
    lib.nsg_ensemble_group_of.argtypes = [vp, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int)]
    lib.nsg_ensemble_n_spect.argtypes = [vp, ctypes.c_int]
    lib.nsg_ensemble_init_state.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_uint64]
    for name in ("nsg_ensemble_set_state", "nsg_ensemble_get_state"):
        getattr(lib, name).argtypes = [vp, ctypes.c_int, vp]
    lib.nsg_ensemble_measures.argtypes = [vp, ctypes.c_int, P(_Totals), vp, vp, ctypes.c_int, ctypes.c_int]
    lib.nsg_ensemble_realspace.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    lib.nsg_ensemble_get_time.argtypes = [vp, ctypes.c_int, dp, dp, dp]
    lib.nsg_ensemble_step.argtypes = [vp, ctypes.c_long]
    lib.nsg_ensemble_step_serial.argtypes = [vp, ctypes.c_long]
    lib.nsg_ensemble_time_steps.argtypes = [vp, ctypes.c_long, ctypes.c_int, dp]
    lib.nsg_ensemble_autotune.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_int)]
    lib.nsg_ensemble_set_schedule.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.nsg_ensemble_step_bytes.argtypes = [vp]
    lib.nsg_ensemble_step_bytes.restype = ctypes.c_double
    # learned closure
    fp = P(ctypes.c_float)
    lib.nsg_cnn_param_count.restype = ctypes.c_longlong
    lib.nsg_cnn_param_count.argtypes = []
    lib.nsg_cnn_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, P(vp)]
    lib.nsg_cnn_destroy.argtypes = [vp]
    lib.nsg_cnn_last_error.argtypes = [vp]
    lib.nsg_cnn_last_error.restype = ctypes.c_char_p
    lib.nsg_cnn_set_params.argtypes = [vp, vp]
    lib.nsg_cnn_get_params.argtypes = [vp, vp]
    lib.nsg_cnn_set_norm.argtypes = [vp, ctypes.c_double, dp]
    lib.nsg_cnn_get_norm.argtypes = [vp, dp, dp]
    lib.nsg_cnn_forward.argtypes = [vp, vp, ctypes.c_int, vp]
    lib.nsg_cnn_loss_grad.argtypes = [vp, vp, vp, ctypes.c_int, dp, vp]
    lib.nsg_cnn_add_sample.argtypes = [vp, vp, vp]
    lib.nsg_cnn_sample_count.argtypes = [vp]
    lib.nsg_cnn_train.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_uint64, dp]
    lib.nsg_cnn_eval.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    lib.nsg_cnn_time_forward.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    lib.nsg_sgs_sample.argtypes = [vp, vp, ctypes.c_double, vp, vp]
    lib.nsg_attach_closure.argtypes = [vp, vp]
    lib.nsg_resample.argtypes = [vp, vp]
    lib.nsg_closure_term.argtypes = [vp, vp]
    lib.nsg_closure_set_gain.argtypes = [vp, vp, ctypes.c_int]
    # increment statistics
    lib.nsg_incr_create.argtypes = [ctypes.c_int, ctypes.c_int, P(vp)]
    lib.nsg_incr_destroy.argtypes = [vp]
    lib.nsg_incr_last_error.argtypes = [vp]
    lib.nsg_incr_last_error.restype = ctypes.c_char_p
    lib.nsg_incr_increments.argtypes = [vp, vp]
    for name in ("nsg_incr_precompute", "nsg_incr_histogram"):
        getattr(lib, name).argtypes = [vp, vp]
    for name in ("nsg_incr_precompute_state", "nsg_incr_histogram_state"):
        getattr(lib, name).argtypes = [vp, vp, ctypes.c_double]
    lib.nsg_incr_set_ranges.argtypes = [vp]
    lib.nsg_incr_results.argtypes = [vp, vp, vp, vp]
    lib.nsg_incr_reset.argtypes = [vp]
    lib.nsg_incr_time_pass.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    _lib = lib
    return lib


class Totals(NamedTuple):
    energy: float
    enstrophy: float
    palinstrophy: float
    helicity: float
    dissipation: float


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def make_params(N: int, nu: float = 1.0, dt: float = 1e-3, visc_pow: int = 1, scheme: str = "rk4cn",
                precision: str = "fp64", cfl: float = 0.0, force_eps: float = 0.0, force_kf: float = 2.5,
                t0: float = 0.0, max_speed: float = 0.0, device: int = 0) -> _Params:
    """An `nsg_params` struct from the library defaults overridden by the keyword values."""
    lib = load_library()
    p = _Params()
    lib.nsg_default_params(ctypes.byref(p))
    p.N[0] = p.N[1] = p.N[2] = int(N)
    p.nu, p.visc_pow, p.dt, p.t0 = float(nu), int(visc_pow), float(dt), float(t0)
    p.scheme, p.precision = SCHEMES[scheme], PRECISIONS[precision]
    p.cfl, p.force_eps, p.force_kf = float(cfl), float(force_eps), float(force_kf)
    p.max_speed, p.device = float(max_speed), int(device)
    return p


def schedule_count(N: int) -> int:
    """Number of launch schedules (FFT schedule x persistent z-pass grid) the library has for size N
    (0 if N is unsupported)."""
    return load_library().nsg_schedule_count(int(N))


class Solver:
    """One periodic box resident on one GPU (a `nsg_ctx`)."""

    def __init__(self, N: int, nu: float = 1.0, dt: float = 1e-3, visc_pow: int = 1, scheme: str = "rk4cn",
                 precision: str = "fp64", cfl: float = 0.0, force_eps: float = 0.0, force_kf: float = 2.5,
                 t0: float = 0.0, max_speed: float = 0.0, device: int = 0):
        self.lib = load_library()
        p = make_params(N, nu=nu, dt=dt, visc_pow=visc_pow, scheme=scheme, precision=precision, cfl=cfl,
                        force_eps=force_eps, force_kf=force_kf, t0=t0, max_speed=max_speed, device=device)
        self.N = int(N)
        self.params = p
        h = ctypes.c_void_p()
        rc = self.lib.nsg_create(ctypes.byref(p), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = self.lib.nsg_last_error(h).decode() if h.value else "nsg_create failed"
            if h.value:
                self.lib.nsg_destroy(h)
                self._h = ctypes.c_void_p()
            raise NSGError(rc, msg)
        self.n_spect = self.lib.nsg_n_spect(h)

    # -- helpers ---------------------------------------------------------------------------
    def _check(self, rc: int):
        if rc != 0:
            raise NSGError(rc, self.lib.nsg_last_error(self._h).decode())

    @property
    def spectral_shape(self):
        return (self.N, self.N, self.N // 2 + 1, 3)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.nsg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state -----------------------------------------------------------------------------
    def set_state(self, u_hat: np.ndarray):
        a = np.ascontiguousarray(u_hat, dtype=np.complex128)
        if a.shape != self.spectral_shape:
            raise ValueError(f"u_hat shape {a.shape} != {self.spectral_shape}")
        self._check(self.lib.nsg_set_state(self._h, _ptr(a)))

    def get_state(self) -> np.ndarray:
        out = np.empty(self.spectral_shape, dtype=np.complex128)
        self._check(self.lib.nsg_get_state(self._h, _ptr(out)))
        return out

    def set_state_real(self, u: np.ndarray):
        a = np.ascontiguousarray(u, dtype=np.float64)
        if a.shape != (self.N,) * 3 + (3,):
            raise ValueError("u must have shape (N, N, N, 3)")
        self._check(self.lib.nsg_set_state_real(self._h, _ptr(a)))

    def init(self, kind: str, k0: float = 4.0, energy: float = 0.5, seed: int = 1):
        self._check(self.lib.nsg_init_state(self._h, ICS[kind], float(k0), float(energy), int(seed)))

    # -- numerics --------------------------------------------------------------------------
    def rhs(self, u_hat: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(u_hat, dtype=np.complex128)
        if a.shape != self.spectral_shape:
            raise ValueError(f"u_hat shape {a.shape} != {self.spectral_shape}")
        out = np.empty(self.spectral_shape, dtype=np.complex128)
        self._check(self.lib.nsg_rhs(self._h, _ptr(a), _ptr(out)))
        return out

    def step(self, nsteps: int = 1):
        """nsteps time steps (scheme "dp5": nsteps accepted adaptive steps)"""
        self._check(self.lib.nsg_step(self._h, int(nsteps)))

    def dp_stats(self):
        """scheme "dp5": (accepted steps, rejected tries, error norm of the last try)"""
        a, r, e = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
        self._check(self.lib.nsg_dp_stats(self._h, ctypes.byref(a), ctypes.byref(r), ctypes.byref(e)))
        return a.value, r.value, e.value

    def measures(self, quirk: bool = False):
        t = _Totals()
        ek = np.zeros(self.n_spect)
        ens = np.zeros(self.n_spect)
        self._check(self.lib.nsg_measures(self._h, ctypes.byref(t), _ptr(ek), _ptr(ens), self.n_spect, int(quirk)))
        return Totals(t.energy, t.enstrophy, t.palinstrophy, t.helicity, t.dissipation), ek, ens

    def realspace(self, which: int = 0) -> np.ndarray:
        out = np.empty((self.N,) * 3 + (3,), dtype=np.float64)
        self._check(self.lib.nsg_realspace(self._h, int(which), _ptr(out)))
        return out

    def time(self):
        t, dt, mx = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.nsg_get_time(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(mx)))
        return t.value, dt.value, mx.value

    def set_time(self, t: float, dt: float):
        self._check(self.lib.nsg_set_time(self._h, float(t), float(dt)))

    def profile(self, iters: int = 3):
        ms = (ctypes.c_double * 6)()
        by = (ctypes.c_double * 5)()
        self._check(self.lib.nsg_profile_passes(self._h, int(iters), ms, by))
        return list(ms), list(by)

    def step_bytes(self) -> float:
        return self.lib.nsg_step_bytes(self._h)

    def set_schedule(self, sched: int):
        self._check(self.lib.nsg_set_schedule(self._h, int(sched)))

    # -- learned closure -------------------------------------------------------------------
    def sgs_sample(self, les: "Solver", kc: float, want_x: bool = True, want_tau: bool = True):
        """Filter this (DNS) box into `les` and return (x [Nl^3, 9], tau [Nl^3, 6]) as float32:
        the velocity gradients of the filtered field and the deviatoric subgrid stress."""
        n = les.N ** 3
        x = np.empty((n, 9), dtype=np.float32) if want_x else None
        t = np.empty((n, 6), dtype=np.float32) if want_tau else None
        rc = self.lib.nsg_sgs_sample(self._h, les._h, float(kc), _ptr(x) if want_x else None,
                                     _ptr(t) if want_tau else None)
        if rc != 0:
            raise NSGError(rc, self.lib.nsg_last_error(les._h).decode() or self.lib.nsg_last_error(self._h).decode())
        return x, t

    def resample_to(self, dst: "Solver"):
        """copy this box's state onto dst's grid (spectral truncation or zero-padding) and time"""
        rc = self.lib.nsg_resample(self._h, dst._h)
        if rc != 0:
            raise NSGError(rc, self.lib.nsg_last_error(dst._h).decode())

    def attach_closure(self, net):
        """Add the network's subgrid term to every later step (None detaches)."""
        self._check(self.lib.nsg_attach_closure(self._h, net._h if net is not None else None))
        self._net = net  # the network must outlive the attachment

    def closure_term(self) -> np.ndarray:
        out = np.empty(self.spectral_shape, dtype=np.complex128)
        self._check(self.lib.nsg_closure_term(self._h, _ptr(out)))
        return out

    def set_closure_gain(self, gain=None):
        """Multiply the attached closure term on the shell b = round(|k|) by gain[b], b = 0..N//3
        (None removes the gains)."""
        if gain is None:
            self._check(self.lib.nsg_closure_set_gain(self._h, None, 0))
            return
        g = np.ascontiguousarray(gain, dtype=np.float64)
        self._check(self.lib.nsg_closure_set_gain(self._h, _ptr(g), int(g.size)))

    def autotune(self, iters: int = 3, use_cache: bool = True):
        """Time every launch schedule of this size on the current state (restored afterwards) and
        keep the fastest; returns (chosen schedule, ms per step of each schedule or None if cached)."""
        n = schedule_count(self.N)
        ch = ctypes.c_int()
        ms = (ctypes.c_double * max(n, 1))(*([-1.0] * max(n, 1)))
        self._check(self.lib.nsg_autotune(self._h, int(iters), int(use_cache), ctypes.byref(ch), ms))
        times = list(ms)[:n]
        return ch.value, (None if all(t < 0 for t in times) else times)


class Ensemble:
    """Many independent boxes on one GPU (a `nsg_ensemble`).  `boxes` is a list of dicts of
    `make_params` keywords, one per box; boxes sharing N, precision, scheme, visc_pow and forcing
    are stepped as one batch, and one step of all batches is a single HIP graph replay."""

    def __init__(self, boxes):
        self.lib = load_library()
        n = len(boxes)
        if n < 1:
            raise ValueError("an ensemble needs at least one box")
        arr = (_Params * n)(*[make_params(**b) for b in boxes])
        self.N = [int(b["N"]) for b in boxes]
        h = ctypes.c_void_p()
        rc = self.lib.nsg_ensemble_create(arr, n, ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = self.lib.nsg_ensemble_last_error(h).decode() if h.value else "nsg_ensemble_create failed"
            if h.value:
                self.lib.nsg_ensemble_destroy(h)
                self._h = ctypes.c_void_p()
            raise NSGError(rc, msg)
        self.nbox = n

    def _check(self, rc: int):
        if rc != 0:
            raise NSGError(rc, self.lib.nsg_ensemble_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.nsg_ensemble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return self.nbox

    @property
    def groups(self) -> int:
        return self.lib.nsg_ensemble_group_count(self._h)

    def group_of(self, box: int):
        g, i = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.nsg_ensemble_group_of(self._h, int(box), ctypes.byref(g), ctypes.byref(i)))
        return g.value, i.value

    def spectral_shape(self, box: int):
        N = self.N[box]
        return (N, N, N // 2 + 1, 3)

    def init(self, box: int, kind: str, k0: float = 4.0, energy: float = 0.5, seed: int = 1):
        self._check(self.lib.nsg_ensemble_init_state(self._h, int(box), ICS[kind], float(k0), float(energy), int(seed)))

    def set_state(self, box: int, u_hat: np.ndarray):
        a = np.ascontiguousarray(u_hat, dtype=np.complex128)
        if a.shape != self.spectral_shape(box):
            raise ValueError(f"u_hat shape {a.shape} != {self.spectral_shape(box)}")
        self._check(self.lib.nsg_ensemble_set_state(self._h, int(box), _ptr(a)))

    def get_state(self, box: int) -> np.ndarray:
        out = np.empty(self.spectral_shape(box), dtype=np.complex128)
        self._check(self.lib.nsg_ensemble_get_state(self._h, int(box), _ptr(out)))
        return out

    def measures(self, box: int, quirk: bool = False):
        ns = self.lib.nsg_ensemble_n_spect(self._h, int(box))
        t = _Totals()
        ek = np.zeros(ns)
        ens = np.zeros(ns)
        self._check(self.lib.nsg_ensemble_measures(self._h, int(box), ctypes.byref(t), _ptr(ek), _ptr(ens), ns,
                                                   int(quirk)))
        return Totals(t.energy, t.enstrophy, t.palinstrophy, t.helicity, t.dissipation), ek, ens

    def realspace(self, box: int, which: int = 0) -> np.ndarray:
        N = self.N[box]
        out = np.empty((N, N, N, 3), dtype=np.float64)
        self._check(self.lib.nsg_ensemble_realspace(self._h, int(box), int(which), _ptr(out)))
        return out

    def time(self, box: int):
        t, dt, mx = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.nsg_ensemble_get_time(self._h, int(box), ctypes.byref(t), ctypes.byref(dt),
                                                   ctypes.byref(mx)))
        return t.value, dt.value, mx.value

    def step(self, nsteps: int = 1):
        self._check(self.lib.nsg_ensemble_step(self._h, int(nsteps)))

    def step_serial(self, nsteps: int = 1):
        self._check(self.lib.nsg_ensemble_step_serial(self._h, int(nsteps)))

    def time_steps(self, nsteps: int, serial: int = 0) -> float:
        """Milliseconds for nsteps steps of every box (HIP events): 0 graph replays, 1 per-box
        launches, 2 per-batch launches on the batches' streams without a graph."""
        ms = ctypes.c_double()
        self._check(self.lib.nsg_ensemble_time_steps(self._h, int(nsteps), int(serial), ctypes.byref(ms)))
        return ms.value

    def autotune(self, iters: int = 3, use_cache: bool = True):
        ch = (ctypes.c_int * max(self.groups, 1))()
        self._check(self.lib.nsg_ensemble_autotune(self._h, int(iters), int(use_cache), ch))
        return list(ch)[:self.groups]

    def set_schedule(self, group: int, sched: int):
        self._check(self.lib.nsg_ensemble_set_schedule(self._h, int(group), int(sched)))

    def step_bytes(self) -> float:
        return self.lib.nsg_ensemble_step_bytes(self._h)


class Network:
    """The learned subgrid-stress network on one GPU (a `nsg_cnn`): periodic 3x3x3 convolutions
    9 -> 32 -> 32 -> 32 -> 6 with GELU, BF16 MFMA kernels, on an N^3 grid."""

    def __init__(self, N: int, max_batch: int = 1, seed: int = 1, device: int = 0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.nsg_cnn_create(int(N), int(max_batch), int(seed), int(device), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = self.lib.nsg_cnn_last_error(h).decode() if h.value else "nsg_cnn_create failed"
            if h.value:
                self.lib.nsg_cnn_destroy(h)
                self._h = ctypes.c_void_p()
            raise NSGError(rc, msg)
        self.N, self.max_batch = int(N), int(max_batch)
        self.param_count = int(self.lib.nsg_cnn_param_count())

    def _check(self, rc: int):
        if rc != 0:
            raise NSGError(rc, self.lib.nsg_cnn_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.nsg_cnn_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def get_params(self) -> np.ndarray:
        out = np.empty(self.param_count, dtype=np.float32)
        self._check(self.lib.nsg_cnn_get_params(self._h, _ptr(out)))
        return out

    def set_params(self, p: np.ndarray):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.param_count:
            raise ValueError(f"expected {self.param_count} parameters")
        self._check(self.lib.nsg_cnn_set_params(self._h, _ptr(a)))

    def set_norm(self, sigma_in: float, sigma_out):
        so = (ctypes.c_double * 6)(*[float(v) for v in sigma_out])
        self._check(self.lib.nsg_cnn_set_norm(self._h, float(sigma_in), so))

    @property
    def norm(self):
        si = ctypes.c_double()
        so = (ctypes.c_double * 6)()
        self._check(self.lib.nsg_cnn_get_norm(self._h, ctypes.byref(si), so))
        return si.value, list(so)

    def _fields(self, a, C):
        a = np.ascontiguousarray(a, dtype=np.float32)
        B = a.size // (self.N ** 3 * C)
        if a.size != B * self.N ** 3 * C or not 1 <= B <= self.max_batch:
            raise ValueError(f"expected [B, N^3, {C}] with 1 <= B <= {self.max_batch}")
        return a, B

    def forward(self, x: np.ndarray) -> np.ndarray:
        """normalised input [B, N^3, 9] (or [B, N, N, N, 9]) -> normalised output [B, N^3, 6]"""
        a, B = self._fields(x, 9)
        out = np.empty((B, self.N ** 3, 6), dtype=np.float32)
        self._check(self.lib.nsg_cnn_forward(self._h, _ptr(a), B, _ptr(out)))
        return out

    def loss_grad(self, x: np.ndarray, t: np.ndarray, grads: bool = True):
        a, B = self._fields(x, 9)
        tt, B2 = self._fields(t, 6)
        if B2 != B:
            raise ValueError("x and t batch sizes differ")
        loss = ctypes.c_double()
        g = np.empty(self.param_count, dtype=np.float32) if grads else None
        self._check(self.lib.nsg_cnn_loss_grad(self._h, _ptr(a), _ptr(tt), B, ctypes.byref(loss),
                                               _ptr(g) if grads else None))
        return loss.value, g

    def add_sample(self, x: np.ndarray, tau: np.ndarray):
        """physical velocity gradients [N^3, 9] and deviatoric stress [N^3, 6]"""
        a, B = self._fields(x, 9)
        t, B2 = self._fields(tau, 6)
        if B != 1 or B2 != 1:
            raise ValueError("one sample at a time")
        self._check(self.lib.nsg_cnn_add_sample(self._h, _ptr(a), _ptr(t)))

    @property
    def samples(self) -> int:
        return self.lib.nsg_cnn_sample_count(self._h)

    def train(self, steps: int, batch: int = 1, lr: float = 1e-3, seed: int = 1, first: int = 0,
              count: int = None) -> np.ndarray:
        """Adam steps on random batches from samples [first, first + count); returns the losses"""
        count = self.samples - first if count is None else count
        hist = np.zeros(max(int(steps), 1))
        self._check(self.lib.nsg_cnn_train(self._h, int(first), int(count), int(steps), int(batch), float(lr),
                                           int(seed), hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return hist[:int(steps)]

    def eval(self, first: int = 0, count: int = None) -> float:
        count = self.samples - first if count is None else count
        loss = ctypes.c_double()
        self._check(self.lib.nsg_cnn_eval(self._h, int(first), int(count), ctypes.byref(loss)))
        return loss.value

    def time_forward(self, batch: int = 1, iters: int = 10) -> float:
        ms = ctypes.c_double()
        self._check(self.lib.nsg_cnn_time_forward(self._h, int(batch), int(iters), ctypes.byref(ms)))
        return ms.value


class IncrementStats:
    """Longitudinal velocity-increment statistics of real-space snapshots on one GPU (the
    reference's post-processing, PostProcessing/src/stats.c): `precompute` every snapshot,
    `set_ranges`, `histogram` every snapshot, then `results()`.  A snapshot is a real AoS field
    (N, N, N, 3) or a Solver (its real-space velocity times `scale`)."""

    NUM, BINS = 9, 1000

    def __init__(self, N: int, device: int = 0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.nsg_incr_create(int(N), int(device), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = self.lib.nsg_incr_last_error(h).decode() if h.value else "nsg_incr_create failed"
            if h.value:
                self.lib.nsg_incr_destroy(h)
                self._h = ctypes.c_void_p()
            raise NSGError(rc, msg)
        self.N = int(N)

    def _check(self, rc: int):
        if rc != 0:
            raise NSGError(rc, self.lib.nsg_incr_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.nsg_incr_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def increments(self) -> np.ndarray:
        r = np.zeros(self.NUM, dtype=np.int32)
        self._check(self.lib.nsg_incr_increments(self._h, _ptr(r)))
        return r

    def _field(self, u):
        a = np.ascontiguousarray(u, dtype=np.float64)
        if a.shape != (self.N,) * 3 + (3,):
            raise ValueError(f"field shape {a.shape} != {(self.N,) * 3 + (3,)}")
        return a

    def precompute(self, snap, scale: float = 1.0):
        if isinstance(snap, Solver):
            self._check(self.lib.nsg_incr_precompute_state(self._h, snap._h, float(scale)))
        else:
            self._check(self.lib.nsg_incr_precompute(self._h, _ptr(self._field(snap))))

    def set_ranges(self):
        self._check(self.lib.nsg_incr_set_ranges(self._h))

    def histogram(self, snap, scale: float = 1.0):
        if isinstance(snap, Solver):
            self._check(self.lib.nsg_incr_histogram_state(self._h, snap._h, float(scale)))
        else:
            self._check(self.lib.nsg_incr_histogram(self._h, _ptr(self._field(snap))))

    def results(self, ranges: bool = True):
        """(stats [9, 6] = min, max, mean, sd, skew, kurtosis; ranges [9, 1001]; counts [9, 1000])"""
        st = np.empty((self.NUM, 6))
        rg = np.empty((self.NUM, self.BINS + 1)) if ranges else None
        ct = np.empty((self.NUM, self.BINS))
        self._check(self.lib.nsg_incr_results(self._h, _ptr(st), _ptr(rg) if ranges else None, _ptr(ct)))
        return st, rg, ct

    def reset(self):
        self._check(self.lib.nsg_incr_reset(self._h))

    def time_pass(self, histogram: bool = False, iters: int = 5) -> float:
        ms = ctypes.c_double()
        self._check(self.lib.nsg_incr_time_pass(self._h, int(histogram), int(iters), ctypes.byref(ms)))
        return ms.value
