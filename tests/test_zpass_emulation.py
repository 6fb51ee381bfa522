"""The persistent z passes (passz_rhs_kernel, passz_rhs2_kernel and passz_full_kernel, the fused
c2r -> u x w -> r2c of NonlinearRHSBatch, Solver/src/solver.c:656-683) run on the CPU through
tests/zpass_emu.cpp, which compiles the device code unchanged and runs each lane as a coroutine that
yields at every wavefront barrier and shuffle (a kernel whose LDS exchanges lacked one would
disagree with the reference here).  Checked line by
line against numpy: c2r_z of the six fields (only kz <= N/3 read, Im X[0] and X[N/2] ignored, as
irfft), the cross product u x w, r2c_z of the three products kept for kz <= N/3 and zero up to the
padded extent, and the running max of |u_x|+|u_y|+|u_z| per box.  Every (precision, N, elements
per thread) that the library builds for N <= 128 is run, several boxes of different amplitudes
per launch and fewer blocks than line groups (so each wavefront loops, crosses box boundaries --
inside one line group too, where N^2 is not a multiple of the lines per wavefront -- and flushes
each box's max), plus the
FP64 N = 16 instance with 8 elements per thread that the library does not build."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

EMU_SRC = os.path.join(ROOT, "tests", "zpass_emu.cpp")
HEADERS = [os.path.join(ROOT, "3d_navier_stokes_amd", "csrc", f) for f in ("nsg_kernels.hpp", "nsg_fft.hpp")]


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("zemu") / "libzemu.so")
    subprocess.run(["g++", "-std=c++20", "-O1", "-fPIC", "-shared", "-I" + os.path.join(ROOT, "tests", "emu"),
                    "-o", so, EMU_SRC], check=True, capture_output=True)
    lib = ctypes.CDLL(so)
    vp, i = ctypes.c_void_p, ctypes.c_int
    lib.emu_passz_rhs.argtypes = [i, i, i, i, i, i, vp, vp, ctypes.c_double, vp]
    lib.emu_passz_rhs.restype = i
    lib.emu_extents.argtypes = [i, vp, vp]
    return lib


def reference(a2, N, kmax, nzp):
    """a2 [nbox][6][N][N][nzp] complex -> ([nbox][3][N][N][nzp], max per box) with unnormalised
    transforms (the kernel's scale is applied to the max only)"""
    X = np.zeros(a2.shape[:-1] + (N // 2 + 1,), dtype=np.complex128)
    X[..., : kmax + 1] = a2[..., : kmax + 1]
    X[..., 0] = X[..., 0].real
    r = np.fft.irfft(X, n=N, axis=-1) * N  # unnormalised c2r
    u, w = r[:, 0:3], r[:, 3:6]
    p = np.stack([u[:, 1] * w[:, 2] - u[:, 2] * w[:, 1], u[:, 2] * w[:, 0] - u[:, 0] * w[:, 2],
                  u[:, 0] * w[:, 1] - u[:, 1] * w[:, 0]], axis=1)
    P = np.fft.rfft(p, axis=-1)
    out = np.zeros(p.shape[:-1] + (nzp,), dtype=np.complex128)
    out[..., : kmax + 1] = P[..., : kmax + 1]
    mx = np.max(np.abs(u).sum(axis=1), axis=(1, 2, 3))
    return out, mx


# (N, elements per thread) of the half-length z transforms of the library's schedules for N <= 128
# (the two radix choices of FFT schedules 0 and 1), per precision; (16, 8) in FP64 is the removed instance
CASES = [("fp64", 8, 4), ("fp64", 16, 4), ("fp64", 16, 8), ("fp64", 32, 8), ("fp64", 32, 4), ("fp64", 64, 8),
         ("fp64", 64, 4), ("fp64", 128, 8), ("fp32", 8, 4), ("fp32", 16, 8), ("fp32", 32, 8), ("fp32", 32, 4),
         ("fp32", 64, 8), ("fp32", 64, 4), ("fp32", 128, 8), ("fp32", 128, 16), ("fp64", 128, 16)]


# kernel versions: 1 passz_rhs_kernel; 2 passz_rhs2_kernel with base twiddles; 3 with the twiddle
# table (elements per thread of the N/2-point transform); 4 passz_full_kernel with base twiddles and
# P_z of two lines per transform, 5 with the twiddle table and P_z alone, 6 with the twiddle table
# and P_z paired (elements per thread of the N-point transform: FULL lists the library's instances
# for N <= 128, including the radix-3 sizes, and one line per wavefront, N/EPT = 64, the geometry
# of the 512^3 instances; N = 12 has no pairs -- 144 lines, 32 per wavefront -- and runs unpaired)
FULL = [("fp64", 8, 4), ("fp32", 8, 4), ("fp64", 12, 6), ("fp32", 12, 6), ("fp64", 16, 4), ("fp32", 16, 4),
        ("fp64", 24, 6), ("fp32", 24, 6), ("fp64", 32, 4), ("fp32", 32, 8), ("fp64", 48, 6), ("fp32", 48, 12),
        ("fp64", 64, 4), ("fp32", 64, 8), ("fp64", 96, 6), ("fp32", 96, 12), ("fp64", 128, 4), ("fp32", 128, 8),
        ("fp64", 128, 2), ("fp64", 64, 8), ("fp32", 16, 2)]
RUNS = [(p, n, e, v) for (p, n, e) in CASES for v in (1, 2, 3)] + [(p, n, e, v) for (p, n, e) in FULL for v in (4, 5, 6)
                                                                   if not (n == 128 and e == 2 and v == 4)]


@pytest.mark.parametrize("prec,N,ept,ver", RUNS)
def test_zpass_matches_numpy(emu, prec, N, ept, ver):
    kmax, nzp = ctypes.c_int(), ctypes.c_int()
    emu.emu_extents(N, ctypes.byref(kmax), ctypes.byref(nzp))
    kmax, nzp = kmax.value, nzp.value
    nbox = 3 if N <= 48 else 1
    rng = np.random.default_rng(N * 31 + ept)
    a2 = (rng.standard_normal((nbox, 6, N, N, nzp)) + 1j * rng.standard_normal((nbox, 6, N, N, nzp))) / N
    # boxes of very different amplitude: a line group that straddles two boxes must not mix their maxima
    a2 *= np.array([1.0, 4.0, 0.25][:nbox])[:, None, None, None, None]
    want, mx = reference(a2, N, kmax, nzp)
    fp32 = prec == "fp32"
    cdt = np.complex64 if fp32 else np.complex128
    inp = np.ascontiguousarray(a2.astype(cdt))
    out = np.full((nbox, 3, N, N, nzp), np.nan, dtype=cdt)  # every element must be written
    got_max = np.zeros(nbox)
    lines_per_wave = 64 // ((N // 2 if ver <= 3 else N) // ept)
    groups = (N * N * nbox) // lines_per_wave
    blocks = max(1, min(3, groups // 4 - 1))  # fewer blocks than groups: persistent loop
    scale = 1.0 / N ** 3
    rc = emu.emu_passz_rhs(N, int(fp32), ept, ver, nbox, blocks, inp.ctypes.data, out.ctypes.data, scale,
                           got_max.ctypes.data)
    assert rc == 0
    assert np.all(np.isfinite(out))
    err = np.linalg.norm(out - want) / np.linalg.norm(want)
    tol = 2e-6 if fp32 else 1e-13
    assert err < tol, err
    # the kz > N/3 pad is exactly zero
    assert np.all(out[..., kmax + 1:] == 0)
    assert np.allclose(got_max, mx * scale, rtol=1e-6 if fp32 else 1e-14, atol=0)
