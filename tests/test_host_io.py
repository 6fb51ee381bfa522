"""Host driver on the CPU: command-line validation (GetCMLArgs, Solver/src/utils.c:73-233) and the
HDF5 layout of the writer (Solver/src/hdf5_funcs.c), checked with synthetic data."""
import os
import subprocess

import numpy as np
import pytest

import h5read
from conftest import ROOT

HOST = os.path.join(ROOT, "3d_navier_stokes_amd", "host")
SOLVER = os.path.join(HOST, "nsg_solver")


@pytest.fixture(scope="module")
def solver_bin():
    subprocess.run(["make", "-C", HOST], check=True, capture_output=True)
    return SOLVER


@pytest.mark.parametrize("args,msg", [(["-n", "63"], "multiple of 2"), (["-n", "2"], "greater than 2"),
                                      (["-h", "0"], "strictly positive"), (["-v", "-1"], "viscosity"),
                                      (["-s", "1.0", "-e", "0.5"], "greater than the initial time"),
                                      (["-c", "0"], "CFL"), (["--scheme", "euler"], "scheme"),
                                      (["-z", "/nonexistent.h5"], "cannot be found")])
def test_cli_rejects_bad_arguments(solver_bin, args, msg):
    p = subprocess.run([solver_bin] + args, capture_output=True, text=True)
    assert p.returncode == 1 and msg in p.stderr


def test_cli_help(solver_bin):
    p = subprocess.run([solver_bin, "--help"], capture_output=True, text=True)
    assert p.returncode == 0 and "-n N" in p.stderr
    assert "rk5cn" in p.stderr and "dp5" in p.stderr


@pytest.mark.parametrize("extra,msg", [(["--adaptive"], "cfl must be 0"), (["--precision", "fp32bulk"], "FP32 round-off")])
def test_cli_dormand_prince_combinations_rejected(solver_bin, tmp_path, extra, msg):
    """the adaptive scheme chooses dt itself and needs FP64: refused before any device call"""
    p = subprocess.run([solver_bin, "-n", "16", "-n", "16", "-n", "16", "--scheme", "dp5", "-o", str(tmp_path) + "/"]
                       + extra, capture_output=True, text=True)
    assert p.returncode == 1 and msg in p.stderr + p.stdout


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("h5")
    exe = str(d / "harness")
    subprocess.run(["gcc", "-O1", "-std=c11", "-D_DEFAULT_SOURCE", f"-I{HOST}", "-I/opt/conda/include",
                    os.path.join(ROOT, "tests", "host_io_harness.c"), os.path.join(HOST, "nsg_io.c"),
                    os.path.join(HOST, "nsg_cli.c"), "-o", exe, "-Wl,-rpath-link,/usr/lib/x86_64-linux-gnu",
                    "/opt/conda/lib/libhdf5_hl.so", "/opt/conda/lib/libhdf5.so", "-Wl,-rpath,/opt/conda/lib", "-lm"],
                   check=True, capture_output=True)
    return exe, d


def test_writer_layout(harness):
    exe, d = harness
    N = 8
    out = subprocess.run([exe, "-n", str(N), "-n", str(N), "-n", str(N), "-o", str(d) + "/", "-v", "0.5", "-e", "2",
                          "-t", "T1", "-i", "TAYLOR_GREEN"], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    main, spect = out.stdout.split()
    assert "SIM_DATA_NAVIER_RK4_FULL_N[8,8]_T[0-2]_NU[0.5000000000]_CFL[1.73]_u0[TAYLOR_GREEN]_TAG[T1]/" in main
    assert main.endswith("Main_HDF_Data.h5") and spect.endswith("Spectra_HDF_Data.h5")
    with h5read.File(main) as f:
        names = f.list("/")
        for n in ["Iter_00000", "Iter_00001", "Iter_00002", "kx", "ky", "kz", "x", "y", "z", "Time", "TotalEnergy",
                  "TotalEnstrophy", "TotalPalinstrophy", "TotalHelicity", "EnergyDissipation"]:
            assert n in names
        assert sorted(f.list("/Iter_00001")) == ["u", "u_hat", "w", "w_hat"]
        assert f.shape("/Iter_00001/u_hat") == (N, N, N // 2 + 1, 3)
        assert f.shape("/Iter_00001/u") == (N, N, N, 3)
        assert f.compound_members("/Iter_00001/u_hat") == ["r", "i"]
        uh = f.read("/Iter_00002/u_hat", "complex")
        flat = np.arange(2 * N * N * (N // 2 + 1) * 3) * 0.5
        assert np.array_equal(uh.reshape(-1), flat[0::2] + 1j * flat[1::2])
        assert f.attr("/Iter_00001", "TimeValue") == 0.5 and f.attr("/Iter_00001", "TimeStep") == 0.125
        assert list(f.read("/kx", "int")) == [0, 1, 2, 3, 4, -3, -2, -1]
        assert list(f.read("/kz", "int")) == [0, 1, 2, 3, 4]
        assert np.allclose(f.read("/x"), 2 * np.pi * np.arange(N) / N)
        assert list(f.read("/Time")) == [0.0, 0.5, 1.0]
    with h5read.File(spect) as f:
        assert sorted(f.list("/Iter_00002")) == ["EnergySpectrum", "EnstrophySpectrum"]
        assert f.shape("/Iter_00002/EnergySpectrum") == (int(np.sqrt(3) * N / 2) + 1,)


def test_file_only_naming(harness):
    exe, d = harness
    out = subprocess.run([exe, "-n", "8", "-n", "8", "-n", "8", "-o", str(d) + "/", "-o", "x"], capture_output=True,
                         text=True)
    assert out.returncode == 0, out.stderr
    assert "Main_HDF_Data_SIM[NAV-RK4-FULL]_N[8,8]_T[0-1]_NU[1.0000000000]_CFL[1.73]_u0[TG_VORT].h5" in out.stdout


CLI_PRINT = r"""
#include <stdio.h>
#include "nsg_run.h"
int main(int argc, char** argv) {
    nsg_run r;
    int rc = nsg_parse_args(argc, argv, &r);
    if (rc) return rc;
    printf("%d %d %d %d %s\n", r.visc_pow, r.transients, r.scheme, r.save_every, r.u0);
    return 0;
}
"""


@pytest.fixture(scope="module")
def cli_print(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    src, exe = str(d / "cli.c"), str(d / "cli")
    with open(src, "w") as f:
        f.write(CLI_PRINT)
    subprocess.run(["gcc", "-O1", "-std=c11", "-D_DEFAULT_SOURCE", f"-I{HOST}", src, os.path.join(HOST, "nsg_cli.c"),
                    "-o", exe, "-lm"], check=True, capture_output=True)
    return exe


def test_cli_defaults_follow_reference_build(cli_print):
    """no flags: the reference's `make solver` physics (hyperviscous, SOLVER_FLAGS -D__HYPER,
    Solver/Makefile:46), RK4, saves every 100 steps (Solver/src/utils.c:45-67), no transients"""
    out = subprocess.run([cli_print], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["2", "0", "0", "100", "TG_VORT"]
    out = subprocess.run([cli_print, "--visc-pow", "1", "--transients", "-i", "BOGUS"], capture_output=True, text=True,
                         check=True).stdout.split()
    assert out == ["1", "1", "0", "100", "NONE"]
