"""End-to-end host driver on the GPU: the Taylor-Green configuration (64^3, nu=1/1600, dt=1e-3,
T=0.2, saves every 10 steps, Solver/src/solver.c:50-208 loop semantics) with both integrators, the
reference's default RANDOM initial condition, the TRANSIENTS build's saving rule, their HDF5 output
against the oracle save by save, and restarts (-z) that continue bitwise or with a new -h."""
import glob
import os
import subprocess

import numpy as np
import pytest

import h5read
from conftest import ROOT, rel

pytestmark = pytest.mark.gpu
SOLVER = os.path.join(ROOT, "3d_navier_stokes_amd", "host", "nsg_solver")
NU = 1.0 / 1600.0


def run(args, tmp):
    p = subprocess.run([SOLVER, "-o", str(tmp) + "/"] + args, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return p.stdout


def main_file(tmp, tag):
    return glob.glob(os.path.join(str(tmp), f"SIM_DATA_*TAG[[]{tag}[]]", "Main_HDF_Data.h5"))[0]


def test_taylor_green_run(oracle, tmp_path):
    out = run(["-n", "64", "-n", "64", "-n", "64", "-s", "0", "-e", "0.2", "-h", "1e-3", "-v", str(NU), "-i",
               "TAYLOR_GREEN", "-p", "10", "-t", "TG", "--no-realspace", "--visc-pow", "1"], tmp_path)
    assert "Total Iters: 200\t Saving Iters: 21" in out
    with h5read.File(main_file(tmp_path, "TG")) as f:
        iters = sorted(n for n in f.list("/") if n.startswith("Iter_"))
        assert iters == [f"Iter_{i:05d}" for i in range(21)]
        t = f.read("/Time")
        assert np.allclose(t, 0.01 * np.arange(21), rtol=0, atol=1e-12)
        E = f.read("/TotalEnergy")
        assert abs(E[0] - np.pi ** 3) < 1e-12 * np.pi ** 3
        assert np.all(np.diff(E) < 0)  # viscous decay, no forcing
        u = oracle.ic_taylor_green(64)
        u = oracle.step(u, NU, 1e-3, nsteps=20)
        assert rel(f.read("/Iter_00002/u_hat", "complex"), u) < 1e-12
        to, _, _ = oracle.measures(u, NU)
        assert abs(E[2] - to[0]) < 1e-12 * to[0]
        assert abs(f.read("/EnergyDissipation")[2] - to[4]) < 1e-12 * to[4]
        assert abs(f.attr("/Iter_00002", "TimeValue") - 0.02) < 1e-15


def test_restart_continues(tmp_path):
    base = ["-n", "32", "-n", "32", "-n", "32", "-h", "1e-3", "-v", "0.01", "-i", "K4", "--k0", "2", "--seed", "5",
            "-p", "5", "--no-realspace", "-f", "CONST_POWER", "--scheme", "ifrk4"]
    run(base + ["-e", "0.02", "-t", "FULL"], tmp_path)
    run(base + ["-e", "0.01", "-t", "HALF"], tmp_path)
    run(base + ["-s", "0.01", "-e", "0.02", "-t", "REST", "-z", main_file(tmp_path, "HALF")], tmp_path)
    with h5read.File(main_file(tmp_path, "FULL")) as a, h5read.File(main_file(tmp_path, "REST")) as b:
        ua = a.read("/Iter_00004/u_hat", "complex")  # t = 0.02
        ub = b.read("/Iter_00002/u_hat", "complex")  # restart at t = 0.01 (Iter_00000) + 10 steps
        assert abs(b.attr("/Iter_00002", "TimeValue") - 0.02) < 1e-15
        assert np.array_equal(ua, ub)


def test_realspace_datasets(oracle, tmp_path):
    run(["-n", "16", "-n", "16", "-n", "16", "-e", "0.002", "-h", "1e-3", "-v", "0.01", "-i", "TAYLOR_GREEN", "-p",
         "1", "-t", "RS"], tmp_path)
    with h5read.File(main_file(tmp_path, "RS")) as f:
        uh = f.read("/Iter_00001/u_hat", "complex")
        assert rel(f.read("/Iter_00001/u"), oracle.realspace(uh, 0)) < 1e-12
        assert rel(f.read("/Iter_00001/w"), oracle.realspace(uh, 1)) < 1e-12
        assert rel(f.read("/Iter_00001/w_hat", "complex"), oracle.curl(uh)) < 1e-14


def test_dormand_prince_run(oracle, tmp_path):
    """--scheme dp5: output named like the reference's -D__DPRK5 build, every save at the accepted
    time of the oracle's adaptive run, states equal to it"""
    run(["-n", "32", "-n", "32", "-n", "32", "-s", "0", "-e", "0.05", "-h", "1e-3", "-v", "0.01", "-i",
         "TAYLOR_GREEN", "-p", "1", "-t", "DP", "--no-realspace", "--scheme", "dp5", "--visc-pow", "1"], tmp_path)
    assert glob.glob(os.path.join(str(tmp_path), "SIM_DATA_NAVIER_DP5_FULL_*TAG[[]DP[]]"))
    with h5read.File(main_file(tmp_path, "DP")) as f:
        t = f.read("/Time")
        u, tt, dt = oracle.ic_taylor_green(32), 0.0, 1e-3
        for i in range(1, 5):
            u, tt, dt, _, _, rc = oracle.dp5_solve(u, 0.01, dt, 1, t=tt)
            assert rc == 0
            assert abs(t[i] - tt) < 1e-9 * tt
            assert abs(f.attr(f"/Iter_{i:05d}", "TimeValue") - tt) < 1e-9 * tt
            assert rel(f.read(f"/Iter_{i:05d}/u_hat", "complex"), u) < 1e-11


def test_taylor_green_ifrk4_every_save(oracle, tmp_path):
    """the Taylor-Green configuration with the integrating-factor scheme: 200 steps, all 21 saves
    (state, totals, E(k)) against the oracle's IF-RK4 trajectory"""
    out = run(["-n", "64", "-n", "64", "-n", "64", "-s", "0", "-e", "0.2", "-h", "1e-3", "-v", str(NU), "-i",
               "TAYLOR_GREEN", "-p", "10", "-t", "TGIF", "--no-realspace", "--scheme", "ifrk4", "--visc-pow", "1"],
              tmp_path)
    assert "Total Iters: 200\t Saving Iters: 21" in out
    u = oracle.ic_taylor_green(64)
    spect = main_file(tmp_path, "TGIF").replace("Main_HDF_Data", "Spectra_HDF_Data")
    with h5read.File(main_file(tmp_path, "TGIF")) as f, h5read.File(spect) as g:
        E, D = f.read("/TotalEnergy"), f.read("/EnergyDissipation")
        for i in range(21):
            if i:
                u = oracle.step(u, NU, 1e-3, 1, "ifrk4", nsteps=10)
            assert rel(f.read(f"/Iter_{i:05d}/u_hat", "complex"), u) < 1e-12, i
            to, eko, _ = oracle.measures(u, NU)
            assert abs(E[i] - to[0]) < 1e-12 * to[0] and abs(D[i] - to[4]) < 1e-12 * to[4], i
            assert rel(g.read(f"/Iter_{i:05d}/EnergySpectrum"), eko) < 1e-12, i


@pytest.mark.parametrize("N", [16, 32])
def test_random_ic_every_save(oracle, tmp_path, N):
    """-i RANDOM, the reference's default branch (Solver/src/solver.c:1605-1624): the glibc
    rand() field after srand(123456789), not solenoidal and not Hermitian, then 2/3 mask; every
    save of a 10-step run against the oracle"""
    out = run(["-n", str(N), "-n", str(N), "-n", str(N), "-s", "0", "-e", "0.01", "-h", "1e-3", "-v", "0.01",
               "--visc-pow", "1", "-i", "RANDOM", "-p", "2", "-t", "RND", "--no-realspace"], tmp_path)
    assert "Using random initial conditions" in out
    u = oracle.ic_random_ref(N)
    with h5read.File(main_file(tmp_path, "RND")) as f:
        assert np.array_equal(f.read("/Iter_00000/u_hat", "complex"), u)
        for i in range(1, 6):
            u2 = oracle.step(u, 0.01, 1e-3, nsteps=2)
            got = f.read(f"/Iter_{i:05d}/u_hat", "complex")
            assert rel(got, u2) < 1e-12, i
            assert rel(got - u, u2 - u) < 1e-11, i  # the increment, nonlinear term included
            u = u2


def test_transients_saving_rule(oracle, tmp_path):
    """--transients (-D__TRANSIENTS): num_print_steps = (num_t_steps - trans)/SAVE_EVERY with
    trans = (long)(0.2 num_t_steps), no initial-condition write, saves when iters > trans and
    iters % SAVE_EVERY == 0, numbered from Iter_00000 (Solver/src/solver.c:120-123,159,1339-1348)"""
    out = run(["-n", "16", "-n", "16", "-n", "16", "-e", "0.05", "-h", "1e-3", "-v", "0.01", "--visc-pow", "1", "-i",
               "K4", "--k0", "2", "--seed", "3", "-p", "5", "-t", "TR", "--no-realspace", "--transients"], tmp_path)
    assert "Total Iters: 50\t Saving Iters: 8\t Transient Steps: 10" in out
    u = oracle.ic_k4(16, 2.0, 0.5, 3)
    u = oracle.step(u, 0.01, 1e-3, nsteps=10)
    with h5read.File(main_file(tmp_path, "TR")) as f:
        iters = sorted(n for n in f.list("/") if n.startswith("Iter_"))
        assert iters == [f"Iter_{i:05d}" for i in range(8)]
        assert np.allclose(f.read("/Time"), 1e-3 * np.arange(15, 51, 5), rtol=0, atol=1e-12)
        for i in range(8):
            u = oracle.step(u, 0.01, 1e-3, nsteps=5)
            assert rel(f.read(f"/Iter_{i:05d}/u_hat", "complex"), u) < 1e-12, i
            assert abs(f.attr(f"/Iter_{i:05d}", "TimeValue") - 1e-3 * (15 + 5 * i)) < 1e-15


def test_restart_with_new_step(oracle, tmp_path):
    """a fixed-dt restart steps with -h even when the file was saved with another TimeStep; saved
    times and states follow the new step"""
    base = ["-n", "16", "-n", "16", "-n", "16", "-v", "0.01", "--visc-pow", "1", "-i", "K4", "--k0", "2", "--seed",
            "4", "-p", "5", "--no-realspace"]
    run(base + ["-h", "1e-3", "-e", "0.01", "-t", "HALF"], tmp_path)
    run(base + ["-h", "5e-4", "-s", "0.01", "-e", "0.02", "-t", "REST", "-z", main_file(tmp_path, "HALF")], tmp_path)
    u = oracle.step(oracle.ic_k4(16, 2.0, 0.5, 4), 0.01, 1e-3, nsteps=10)
    with h5read.File(main_file(tmp_path, "REST")) as f:
        u0 = f.read("/Iter_00000/u_hat", "complex")  # the HALF run's last save
        assert rel(u0, u) < 1e-12
        assert abs(f.attr("/Iter_00000", "TimeStep") - 5e-4) < 1e-18
        u = u0
        for i in range(1, 5):
            u = oracle.step(u, 0.01, 5e-4, nsteps=5)
            assert abs(f.attr(f"/Iter_{i:05d}", "TimeValue") - (0.01 + 0.0025 * i)) < 1e-14, i
            assert abs(f.attr(f"/Iter_{i:05d}", "TimeStep") - 5e-4) < 1e-18, i
            assert rel(f.read(f"/Iter_{i:05d}/u_hat", "complex"), u) < 1e-12, i
