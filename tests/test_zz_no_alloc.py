"""nsg_measures_noalloc (the allocation-free variant of nsg_measures) through the C ABI, against
nsg_measures on the same state: totals and both spectra bit for bit, for two calls in a row on one
GPU -- the second on a smaller grid, so it runs on a static scratch array that still holds the
first call's larger set of partial sums -- with and without the reference's quirk arithmetic.
The output buffer is device memory allocated here (by the caller) with the HIP runtime the library
links against; the entry point itself allocates nothing."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime libnsg.so is linked against (already loaded)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return hip


HIP_D2H = 2


def _noalloc(nsg, s, hip, quirk):
    ns = s.n_spect
    n = 5 + 2 * ns
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), 8 * n) == 0
    try:
        assert hip.hipMemset(d, 0xFF, 8 * n) == 0  # NaN pattern: every element must be written
        rc = s.lib.nsg_measures_noalloc(s._h, d, ns, int(quirk))
        assert rc == 0, s.lib.nsg_last_error(s._h)
        assert hip.hipDeviceSynchronize() == 0
        host = np.empty(n)
        assert hip.hipMemcpy(host.ctypes.data, d, 8 * n, HIP_D2H) == 0
    finally:
        hip.hipFree(d)
    return host[:5], host[5:5 + ns], host[5 + ns:]


@pytest.mark.parametrize("quirk", [False, True])
def test_measures_noalloc_matches_measures(nsg, quirk):
    hip = _hip()
    for N, seed in ((64, 3), (32, 4)):  # second call: fewer spectral shells than the first
        with nsg.Solver(N, nu=0.01, visc_pow=1) as s:
            s.init("k4", k0=2.0, energy=0.5, seed=seed)
            s.step(2)
            tot, ek, ens = s.measures(quirk=quirk)
            t2, e2, n2 = _noalloc(nsg, s, hip, quirk)
            assert np.array_equal(t2, np.array([tot.energy, tot.enstrophy, tot.palinstrophy, tot.helicity,
                                                tot.dissipation]))
            assert np.array_equal(e2, ek) and np.array_equal(n2, ens)


def test_measures_noalloc_rejects_short_output(nsg):
    with nsg.Solver(16) as s:
        rc = s.lib.nsg_measures_noalloc(s._h, ctypes.c_void_p(16), s.n_spect - 1, 0)  # never dereferenced
        assert rc == -1 and b"n_spect" in s.lib.nsg_last_error(s._h)
