"""CPU AddressSanitizer + UndefinedBehaviorSanitizer builds of the C code that runs on the host:
the oracle (oracle/nsg_oracle.c, oracle/stats_oracle.c, driven by tests/sanitize_driver.c) and the
host driver's command line and HDF5 writer/restart reader (host/nsg_cli.c, host/nsg_io.c, driven by
tests/host_io_harness.c).  Any report aborts the program (halt_on_error / -fno-sanitize-recover),
so a clean exit means no out-of-bounds access, leak or undefined arithmetic on these paths.  The
device code is not built here (GPU sanitizers are not available for it)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HOST = os.path.join(ROOT, "3d_navier_stokes_amd", "host")
ORACLE = os.path.join(ROOT, "oracle")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_driver")
    subprocess.run(["gcc", "-std=gnu11"] + SAN + [f"-I{ORACLE}", os.path.join(ROOT, "tests", "sanitize_driver.c"),
                    os.path.join(ORACLE, "nsg_oracle.c"), os.path.join(ORACLE, "stats_oracle.c"), "-o", exe, "-lm"],
                   check=True, capture_output=True)
    p = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=600)
    assert p.returncode == 0 and p.stdout.strip() == "ok", p.stderr[-4000:]


def test_host_io_under_sanitizers(tmp_path):
    exe = str(tmp_path / "harness")
    subprocess.run(["gcc", "-std=c11", "-D_DEFAULT_SOURCE"] + SAN + [f"-I{HOST}", "-I/opt/conda/include",
                    os.path.join(ROOT, "tests", "host_io_harness.c"), os.path.join(HOST, "nsg_io.c"),
                    os.path.join(HOST, "nsg_cli.c"), "-o", exe, "-Wl,-rpath-link,/usr/lib/x86_64-linux-gnu",
                    "/opt/conda/lib/libhdf5_hl.so", "/opt/conda/lib/libhdf5.so", "-Wl,-rpath,/opt/conda/lib", "-lm"],
                   check=True, capture_output=True)
    out = str(tmp_path) + "/"
    # HDF5's own one-time allocations are process-lifetime: leak checking covers this code only
    env = dict(ENV, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    for extra in ([], ["-o", out]):  # directory output and file-only naming
        p = subprocess.run([exe, "-n", "8", "-n", "8", "-n", "8", "-o", out, "-t", "SAN", "--transients"] + extra,
                           capture_output=True, text=True, env=env, timeout=300)
        assert p.returncode == 0, p.stderr[-4000:]
    for bad in (["-n", "7"], ["--scheme", "euler"], ["--visc-pow", "9"]):
        p = subprocess.run([exe] + bad, capture_output=True, text=True, env=env, timeout=60)
        assert p.returncode == 3 and "ERROR" in p.stderr, p.stderr[-2000:]
