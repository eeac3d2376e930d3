"""The C-ABI library loads without a GPU and exports every entry point that
include/snapx.h declares; the ctypes binding covers all of them."""
import ctypes
import os
import re

from snapy_linfel_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "snapx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(snapx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "snapx_hydro_forward" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(capi.SIGNATURES) == declared_functions()


def test_version_and_no_device_needed():
    L = capi.lib()
    assert L.snapx_version() == 1
    assert L.snapx_device_count() >= 0


def test_config_layout_matches_oracle():
    import oracle as O
    assert ctypes.sizeof(capi.Config) == ctypes.sizeof(O.OcConfig)
    assert [f[0] for f in capi.Config._fields_] == [f[0] for f in O.OcConfig._fields_]


def test_workspace_queries_without_device():
    c = capi.make_config(nx1=64, nx2=512, nghost=3, implicit_scheme=1)
    L = capi.lib()
    assert L.snapx_hydro_workspace_bytes(ctypes.byref(c), capi.F64) > 0
    assert L.snapx_implicit_workspace_bytes(ctypes.byref(c)) == 8 * (64 * 512 * (27 + 3) + 2 * 512)
    bad = capi.make_config(nx1=2, nghost=3)
    assert L.snapx_hydro_workspace_bytes(ctypes.byref(bad), capi.F64) == 0
