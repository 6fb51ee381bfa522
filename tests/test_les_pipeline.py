"""CPU checks of the LES pipeline's host-side statistics (3d_navier_stokes_amd/les.py): the per-shell
subgrid transfer spectrum behind the per-shell closure calibration."""
import importlib.util
import os

import numpy as np

from conftest import ROOT

_spec = importlib.util.spec_from_file_location("les", os.path.join(ROOT, "3d_navier_stokes_amd", "les.py"))
les = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(les)


def _band_limited(rng, N, nch, kmax):
    """real fields [N^3, nch] whose modes all lie on shells round(|k|) <= kmax"""
    k = np.fft.fftfreq(N, 1.0 / N)
    kz = np.arange(N // 2 + 1)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
    keep = np.sqrt(k2) < kmax + 0.5
    out = np.empty((N, N, N, nch))
    for c in range(nch):
        h = (rng.standard_normal(k2.shape) + 1j * rng.standard_normal(k2.shape)) * keep
        out[..., c] = np.fft.irfftn(h, s=(N, N, N), axes=(0, 1, 2))
    return out.reshape(-1, nch)


def test_transfer_parseval():
    """a gradient field inside the 2/3 sphere: the shells' transfers add up to N^3 sum_x S_ij tau_ij"""
    N = 24
    rng = np.random.default_rng(4)
    x = _band_limited(rng, N, 9, N // 3)
    tau = rng.standard_normal((N ** 3, 6))
    T = les.sgs_transfer(x, tau, N)
    g = x.reshape(-1, 3, 3)
    S = 0.5 * (g + g.transpose(0, 2, 1))
    direct = np.sum(tau[:, 0] * S[:, 0, 0] + tau[:, 1] * S[:, 1, 1] + tau[:, 2] * S[:, 2, 2]
                    + 2 * (tau[:, 3] * S[:, 0, 1] + tau[:, 4] * S[:, 0, 2] + tau[:, 5] * S[:, 1, 2]))
    assert T.shape == (N // 3 + 1,)
    assert abs(T.sum() - N ** 3 * direct) < 1e-9 * abs(N ** 3 * direct)


def test_transfer_single_shell():
    """S and tau of one Fourier mode k = (2, 1, 2), |k| = 3: all of the transfer lands on shell 3, with
    the value of the one mode (kz > 0: counted twice in the half spectrum)"""
    N = 16
    i, j, l = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    ph = 2 * np.pi * (2 * i + 1 * j + 2 * l) / N
    x = np.zeros((N, N, N, 9))
    x[..., 1] = np.cos(ph)  # d_y u_x: S_xy = S_yx = cos / 2
    tau = np.zeros((N, N, N, 6))
    tau[..., 3] = 3.0 * np.cos(ph)  # tau_xy
    T = les.sgs_transfer(x.reshape(-1, 9), tau.reshape(-1, 6), N)
    want = N ** 3 * np.sum(2 * 0.5 * np.cos(ph) * 3.0 * np.cos(ph))
    assert np.allclose(T[[0, 1, 2, 4, 5]], 0.0, atol=1e-9 * abs(want))
    assert abs(T[3] - want) < 1e-9 * abs(want)
