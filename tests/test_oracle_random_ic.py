"""Pin of the reference's default RANDOM initial condition (Solver/src/solver.c:1551,1605-1624).

The reference seeds glibc's generator with srand(123456789) and, looping i (x), j (y), k (kz) with
the component d innermost, sets u_hat = U1 * cexp(2 pi i U2) with U = rand() / RAND_MAX; then
ApplyDealiasing zeroes |k| > floor(N/3).  Two facts pin the oracle's restatement here:

* the glibc stream itself, drawn through ctypes from this process's libc (the same generator the
  reference links), assembled by numpy in (i, j, k, d) order and masked by an independent numpy
  2/3 rule -- this fixes the loop order, the component order, the scaling of rand() and the mask;
* the order of the two rand() calls inside the reference's one expression, which C leaves
  unspecified: gcc (the compiler of the reference's Makefile, mpicc, Solver/Makefile:15-17)
  is asked directly by compiling an expression of the same shape at -O0 and -O3.
"""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import wavenumbers

SEED = 123456789
RAND_MAX = 2147483647  # glibc


def _glibc_stream(n):
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(ctypes.c_uint(SEED))
    return [libc.rand() for _ in range(n)]


def _sincos():
    libm = ctypes.CDLL("libm.so.6")
    f = libm.sincos
    f.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    f.restype = None
    s, c = ctypes.c_double(), ctypes.c_double()

    def sc(x):
        f(x, ctypes.byref(s), ctypes.byref(c))
        return s.value, c.value
    return sc


def _field(N, amplitude_first):
    nz = N // 2 + 1
    n = N * N * nz * 3
    r = _glibc_stream(2 * n)
    sincos = _sincos()
    out = np.empty(n, dtype=np.complex128)
    for m in range(n):
        a = float(r[2 * m]) / float(RAND_MAX)
        b = float(r[2 * m + 1]) / float(RAND_MAX)
        amp, ph = (a, b) if amplitude_first else (b, a)
        # cexp(i y) in glibc is (cos y, sin y) from sincos(y) (checked by the probe below)
        sn, cs = sincos(ph * 2.0 * math.pi)
        out[m] = complex(amp * cs, amp * sn)
    out = out.reshape(N, N, nz, 3)
    KX, KY, KZ = wavenumbers(N)
    keep = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2) <= N // 3
    return np.where(keep[..., None], out, 0)


@pytest.mark.parametrize("N", [8, 12, 16])
def test_random_ic_matches_glibc_stream(oracle, N):
    got = oracle.ic_random_ref(N)
    want = _field(N, amplitude_first=True)
    assert np.array_equal(got, want)
    other = _field(N, amplitude_first=False)
    assert not np.array_equal(got, other)
    # not solenoidal and not Hermitian, as in the reference: the state is used as drawn
    assert np.count_nonzero(got) == 3 * np.count_nonzero(want[..., 0])


PROBE = r"""
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
int main(void) {
    srand(123456789);
    /* the reference's expression shape: amplitude factor times cexp of the phase factor */
    double complex z = ((double)rand() / (double)RAND_MAX) * cexp(((double)rand() / (double)RAND_MAX) * 2.0 * M_PI * I);
    srand(123456789);
    double first = (double)rand() / (double)RAND_MAX;
    double second = (double)rand() / (double)RAND_MAX;
    double complex amp_first = first * cexp(second * 2.0 * M_PI * I);
    double complex phase_first = second * cexp(first * 2.0 * M_PI * I);
    /* cexp of a purely imaginary argument equals glibc's sincos, bit for bit, over many draws */
    int same = 1;
    for (int n = 0; n < 200000; ++n) {
        double y = (double)rand() / (double)RAND_MAX * 2.0 * M_PI, s, c;
        double complex e = cexp(y * I);
        sincos(y, &s, &c);
        same &= (creal(e) == c) && (cimag(e) == s);
    }
    printf("%d %d %d\n", z == amp_first, z == phase_first, same);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
@pytest.mark.parametrize("opt", ["-O0", "-O3"])
def test_gcc_draws_amplitude_first(tmp_path, opt):
    """gcc evaluates the left factor's rand() first, so the amplitude is the first draw -- the order
    the oracle and the host driver use -- and glibc's cexp(i y) is sincos(y) bit for bit."""
    src = os.path.join(tmp_path, "order.c")
    exe = os.path.join(tmp_path, "order")
    with open(src, "w") as f:
        f.write(PROBE)
    subprocess.run(["gcc", opt, "-std=gnu11", "-o", exe, src, "-lm"], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    assert out == ["1", "0", "1"]
