"""The C-ABI library loads without a GPU and exports every function that
include/channel_flow.h declares (no compute calls here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    txt = (ROOT / "include" / "channel_flow.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:[a-z_]+\s+)*[a-z_]+\s*\*?\s*(cf_[a-z0-9_]+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("cf_create", "cf_flux_x", "cf_flux_y", "cf_flux_z", "cf_viscous_terms", "cf_dual_time_stepping",
              "cf_inner_iteration", "cf_physical_step", "cf_write_output", "cf_jacobian_all", "cf_hyperplane"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib_path = ROOT / "channel_flow_amd" / "libchannel_flow.so"
    if not lib_path.exists():
        pytest.fail(f"{lib_path} not built (python -m channel_flow_amd.build)")
    lib = ctypes.CDLL(str(lib_path))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_export():
    from channel_flow_amd.solver import EXPORTS
    assert sorted(EXPORTS) == declared_functions()


def declared_arity():
    """name -> number of parameters, from the prototypes in the header"""
    txt = (ROOT / "include" / "channel_flow.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for name, params in re.findall(r"(cf_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", txt, flags=re.S):
        params = params.strip()
        out[name] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_python_mirror_declares_every_signature():
    """Each bound export carries argtypes of the header's arity, so ctypes
    converts doubles and pointers instead of guessing."""
    from channel_flow_amd.solver import EXPORTS, load_library
    lib = load_library()
    arity = declared_arity()
    for n in EXPORTS:
        at = getattr(lib, n).argtypes
        assert at is not None, n
        assert len(at) == arity[n], (n, len(at), arity[n])


def test_solver_refuses_without_library(tmp_path, monkeypatch):
    """The product path fails loudly instead of falling back to anything."""
    import channel_flow_amd.solver as S
    monkeypatch.setattr(S, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(S, "_lib", None)
    with pytest.raises(S.CfError):
        S.load_library()


def test_create_rejects_bad_params():
    """cf_create validates sizes before touching the device."""
    from channel_flow_amd.solver import CfParams, load_library
    lib = load_library()
    h = ctypes.c_void_p()
    p = CfParams(4, 64, 64, 1e-3, 1 / 64, 1e-3, 0.005, 600.0, 5e-8, 0, 2, 200, 1e-3, 0)
    assert lib.cf_create(ctypes.byref(p), ctypes.byref(h)) == -1
    p = CfParams(64, 64, 64, 1e-3, 1 / 64, 1e-3, 0.005, 600.0, -1.0, 0, 2, 200, 1e-3, 0)
    assert lib.cf_create(ctypes.byref(p), ctypes.byref(h)) == -1
