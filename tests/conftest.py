"""Shared test fixtures.  Tests marked `gpu` need an MI355X; everything else runs on the CPU."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(ROOT, "tests")
GOLDEN = os.path.join(TESTS, "golden")
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X) and the built libnsg.so")


def load_nsg():
    pkg = os.path.join(ROOT, "3d_navier_stokes_amd")
    if "navier_stokes_amd" in sys.modules:
        return sys.modules["navier_stokes_amd"]
    spec = importlib.util.spec_from_file_location("navier_stokes_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["navier_stokes_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def nsg():
    return load_nsg()


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib

    oracle_lib.load()
    return oracle_lib


@pytest.fixture(scope="session")
def golden():
    return {"fft": np.load(os.path.join(GOLDEN, "fft.npz")), "phys": np.load(os.path.join(GOLDEN, "physics.npz"))}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def wavenumbers(N):
    k = np.array([i if i <= N // 2 else i - N for i in range(N)])
    kz = np.arange(N // 2 + 1)
    return np.meshgrid(k, k, kz, indexing="ij")
