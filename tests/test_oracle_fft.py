"""The oracle's own FFT against numpy (pocketfft) golden vectors, brute-force DFTs and identities."""
import numpy as np
import pytest

from conftest import rel


def brute_dft(x, sign):
    n = x.shape[0]
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * (np.outer(k, k) % n) / n) @ x  # exponents reduced mod n


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 9, 12, 16, 24, 27, 32, 48, 96])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_brute_force(oracle, n, sign):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    assert rel(oracle.fft1d(x, sign), brute_dft(x, sign)) < 1e-14


@pytest.mark.parametrize("N", [8, 12, 16, 24])
def test_r2c3d_golden(oracle, golden, N):
    g = golden["fft"]
    assert rel(oracle.r2c3d(g[f"fft_in_{N}"]), g[f"fft_out_{N}"]) < 1e-14


@pytest.mark.parametrize("N", [8, 12, 16, 24])
def test_c2r3d_golden_roundtrip(oracle, golden, N):
    g = golden["fft"]
    back = oracle.c2r3d(g[f"fft_out_{N}"]) / N ** 3
    assert np.max(np.abs(back - g[f"fft_in_{N}"])) < 1e-13


def test_r2c3d_brute_force_small(oracle):
    N = 6
    x = np.random.default_rng(3).standard_normal((N, N, N))
    F = np.exp(-2j * np.pi * np.outer(np.arange(N), np.arange(N)) / N)
    full = np.einsum("ai,bj,ck,ijk->abc", F, F, F, x)
    assert rel(oracle.r2c3d(x), full[:, :, : N // 2 + 1]) < 1e-14


def test_c2r_uses_real_part_of_selfconjugate_lines(oracle):
    """c2r drops Im of the kz=0 and kz=N/2 coefficients, as numpy.fft.irfftn does."""
    N = 8
    X = np.random.default_rng(5).standard_normal((N, N, N // 2 + 1)) * (1 + 1j)
    assert rel(oracle.c2r3d(X), np.fft.irfftn(X, s=(N, N, N), axes=(0, 1, 2)) * N ** 3) < 1e-14


@pytest.mark.parametrize("N", [12, 16])
def test_parseval(oracle, N):
    x = np.random.default_rng(N).standard_normal((N, N, N))
    X = oracle.r2c3d(x)
    w = np.ones(N // 2 + 1) * 2.0
    w[0] = 1.0
    w[-1] = 1.0
    assert abs(np.sum(w * np.abs(X) ** 2) / N ** 3 - np.sum(x * x)) < 1e-10 * np.sum(x * x)
