"""libnsg.so loads and exports every function declared in include/nsg.h; argument validation that
fails before any device call.  No compute call is made here (no GPU in the CPU suite)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "nsg.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nsg_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ["nsg_create", "nsg_destroy", "nsg_set_state", "nsg_get_state", "nsg_step", "nsg_rhs",
                     "nsg_measures", "nsg_realspace", "nsg_last_error"]:
        assert required in names


def test_library_exports_every_declared_symbol(nsg):
    lib = nsg.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_default_params_mirror_reference_cli_defaults(nsg):
    """Solver/src/utils.c:45-67: N = 64^3, dt = 1e-3, nu = 1.0"""
    lib = nsg.load_library()
    p = nsg._Params()
    lib.nsg_default_params(ctypes.byref(p))
    assert list(p.N) == [64, 64, 64] and p.dt == 1e-3 and p.nu == 1.0 and p.scheme == 0 and p.precision == 0


@pytest.mark.parametrize("kw,msg", [({"N": 100}, "unsupported N"), ({"N": 64, "dt": 0.0}, "dt"),
                                    ({"N": 64, "visc_pow": 7}, "visc_pow"),
                                    ({"N": 64, "force_eps": 0.1, "force_kf": 9.0}, "force_kf"),
                                    ({"N": 64, "scheme": "dp5", "cfl": 0.5}, "cfl must be 0"),
                                    ({"N": 64, "scheme": "dp5", "precision": "fp32bulk"}, "FP32 round-off")])
def test_invalid_parameters_are_rejected_before_device_use(nsg, kw, msg):
    N = kw.pop("N")
    with pytest.raises(nsg.NSGError) as e:
        nsg.Solver(N, **kw)
    assert e.value.code == -1 and msg in str(e.value)


def test_non_cubic_grid_rejected(nsg):
    lib = nsg.load_library()
    p = nsg._Params()
    lib.nsg_default_params(ctypes.byref(p))
    p.N[2] = 32
    h = ctypes.c_void_p()
    assert lib.nsg_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"cubic" in lib.nsg_last_error(h)
    lib.nsg_destroy(h)


def test_dormand_prince_is_single_box(nsg):
    """the adaptive scheme accepts or rejects each box's try on the host: no ensembles of it"""
    with pytest.raises(nsg.NSGError) as e:
        nsg.Ensemble([dict(N=16, scheme="dp5"), dict(N=16, scheme="dp5")])
    assert e.value.code == -1 and "one box per context" in str(e.value)
