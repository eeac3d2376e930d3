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


def test_spherical_polar_limits_without_device():
    """Spherical-polar blocks run on the plain and fused paths (fused columns
    up to one column workgroup) and in the adjoint (columns up to one column
    workgroup); every path refuses blocks without a radial axis or with
    theta outside [0, pi] (src/coord/spherical_polar.cpp:46-48).  Gnomonic
    panels: the same fused column limit; the adjoint refuses them."""
    L = capi.lib()
    L.snapx_last_error.restype = ctypes.c_char_p
    kw = dict(nx1=16, nx2=8, nx3=8, nghost=3, x1min=1.0, x1max=2.0, x2min=0.7, x2max=2.4,
              x3min=0.0, x3max=1.0, coord="spherical-polar", implicit_scheme=1)
    c = capi.make_config(**kw)
    assert L.snapx_hydro_workspace_bytes(ctypes.byref(c), capi.F64) > 0
    assert L.snapx_rk3_workspace_bytes(ctypes.byref(c), capi.F64) > 0
    assert L.snapx_fused_workspace_bytes(ctypes.byref(c)) > 0
    tall = capi.make_config(**dict(kw, nx1=300))
    assert L.snapx_fused_workspace_bytes(ctypes.byref(tall)) == 0
    assert b"curvilinear columns of at most" in L.snapx_last_error()
    assert L.snapx_adjoint_workspace_bytes(ctypes.byref(c)) > 0
    assert L.snapx_adjoint_f32_workspace_bytes(ctypes.byref(c)) > 0
    assert L.snapx_adjoint_workspace_bytes(ctypes.byref(tall)) == 0
    assert b"spherical-polar columns of at most" in L.snapx_last_error()
    pi4 = 0.785
    gk = dict(kw, nx2=10, nx3=10, x2min=-pi4, x2max=pi4, x3min=-pi4, x3max=pi4,
              coord="gnomonic-equiangle")
    assert L.snapx_fused_workspace_bytes(ctypes.byref(capi.make_config(**gk))) > 0
    assert L.snapx_fused_workspace_bytes(ctypes.byref(capi.make_config(**dict(gk, nx1=300)))) == 0
    assert b"curvilinear columns of at most" in L.snapx_last_error()
    assert L.snapx_adjoint_workspace_bytes(ctypes.byref(capi.make_config(**gk))) == 0
    assert b"Cartesian and spherical-polar" in L.snapx_last_error()
    for bad in (dict(kw, x2max=3.3), dict(kw, nx1=1)):
        assert L.snapx_hydro_workspace_bytes(ctypes.byref(capi.make_config(**bad)), capi.F64) == 0
    c.coord = 7
    assert L.snapx_hydro_workspace_bytes(ctypes.byref(c), capi.F64) == 0
    assert b"Unknown coordinate type" in L.snapx_last_error()
