"""ctypes wrapper of the CPU oracle (oracle/snap_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker / CPU timing baseline.
The shipped package never imports this module.

Arrays are float64 numpy arrays in the reference layout ``[nvar, n3, n2, n1]``
(x1 fastest).  ``dim`` is the reference tensor axis: 3 = x1, 2 = x2, 1 = x3.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

IDN, IVX, IVY, IVZ, IPR, ICY = 0, 1, 2, 3, 4, 5
ILT, IRT = 0, 1
INTERP = {"dc": 0, "plm": 1, "cp3": 2, "cp5": 3, "weno3": 4, "weno5": 5}
RIEMANN = {"roe": 0, "lmars": 1, "hllc": 2}
BC = {None: -1, "none": -1, "reflecting": 0, "outflow": 1, "periodic": 2, "custom": 3}
COORD = {"cartesian": 0, "spherical-polar": 1, "gnomonic-equiangle": 2}
PRIMITIVE, CONSERVED, SCALAR = 0, 1, 2


class OcConfig(C.Structure):
    _fields_ = [
        ("nx1", C.c_int), ("nx2", C.c_int), ("nx3", C.c_int), ("nghost", C.c_int),
        ("x1min", C.c_double), ("x1max", C.c_double),
        ("x2min", C.c_double), ("x2max", C.c_double),
        ("x3min", C.c_double), ("x3max", C.c_double),
        ("gammad", C.c_double),
        ("recon_type", C.c_int * 2), ("recon_scale", C.c_int * 2),
        ("recon_shock", C.c_int * 2),
        ("riemann", C.c_int * 3),
        ("limiter", C.c_int),
        ("density_floor", C.c_double), ("pressure_floor", C.c_double),
        ("temperature_floor", C.c_double), ("weight", C.c_double),
        ("grav1", C.c_double), ("grav2", C.c_double), ("grav3", C.c_double),
        ("non_hydrostatic", C.c_double),
        ("implicit_scheme", C.c_int),
        ("disable_flux", C.c_int * 3),
        ("bc", C.c_int * 6),
        ("coord", C.c_int),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(C.c_double)
        cfgp = C.POINTER(OcConfig)
        d = C.c_double
        i = C.c_int
        for name in ("oc_interp_cp3", "oc_interp_weno3", "oc_interp_plm"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [d, d, d]
        for name in ("oc_interp_cp5", "oc_interp_weno5"):
            getattr(L, name).restype = d
            getattr(L, name).argtypes = [d, d, d, d, d]
        L.oc_interp_window.restype = d
        L.oc_interp_window.argtypes = [i, i, i, P]
        L.oc_cons2prim.argtypes = [cfgp, P, P, C.c_int64]
        L.oc_prim2cons.argtypes = [cfgp, P, P, C.c_int64]
        L.oc_gno_prim2local.argtypes = [cfgp, P, i]
        L.oc_gno_flux2global.argtypes = [cfgp, P, i]
        L.oc_reconstruct.argtypes = [P, P, i, i, i, i, i, i, i, i, i, i, d, d]
        L.oc_riemann.argtypes = [i, P, P, P, P, i, i, i, i, i, d]
        L.oc_divergence.argtypes = [P, P, P, P, i, i, i, i, i, P, P, P]
        L.oc_coord_forward.argtypes = [cfgp, P, P, P, P, P, P]
        L.oc_geometry.argtypes = [cfgp, i, P]
        L.oc_revise_x1_ghost.argtypes = [P, i, i, i, i, i, d, d, P]
        L.oc_revise_x1_lr.argtypes = [P, i, i, i, i, i]
        L.oc_implicit_correction.argtypes = [cfgp, d, P, P]
        L.oc_vic_backward_reduce.argtypes = [i, P, P, P, P, P, P, i, i, i, i, i, i]
        L.oc_vic_redistribute_cell.argtypes = [i, P, P, P, P, P, P, i, i, i, i, i]
        L.oc_inverse3.argtypes = [P, P]
        L.oc_vic_column_matrices.argtypes = [cfgp, d, P, i, i, P, P, P]
        L.oc_flux_jacobian5.argtypes = [d, P, P]
        L.oc_roe_state.argtypes = [d, P, P, P, P]
        L.oc_roe_eigensystem.argtypes = [d, P, P, P, P, P, P]
        L.oc_vic_assemble.argtypes = [i, P, C.c_long, i, d, P, P, d, d, P, P, P]
        L.oc_hydro_forward.argtypes = [cfgp, d, P, P, P, P, P, P]
        L.oc_apply_bc.argtypes = [cfgp, P, i, i]
        L.oc_max_time_step.restype = d
        L.oc_max_time_step.argtypes = [cfgp, P, d]
        L.oc_rk3_step.argtypes = [cfgp, d, P, P]
        L.oc_initialize.argtypes = [cfgp, P, P]
        L.oc_set_threads.argtypes = [i]
        L.oc_get_threads.restype = i
        _lib = L
    return _lib


def _check_field(c, a, nvar=None):
    """Shape guard: the C oracle trusts its sizes, so every field passed with a
    config must be [nvar, n3, n2, n1] of that config."""
    assert a.ndim == 4 and tuple(a.shape[1:]) == tuple(shape(c)), (a.shape, shape(c))
    if nvar is not None:
        assert a.shape[0] == nvar, (a.shape, nvar)


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous"
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_config(**kw) -> OcConfig:
    """Build an OcConfig from keyword options (names follow the YAML keys)."""
    c = OcConfig()
    c.nx1 = kw.get("nx1", 1)
    c.nx2 = kw.get("nx2", 1)
    c.nx3 = kw.get("nx3", 1)
    c.nghost = kw.get("nghost", 3)
    c.x1min, c.x1max = kw.get("x1min", 0.0), kw.get("x1max", 1.0)
    c.x2min, c.x2max = kw.get("x2min", 0.0), kw.get("x2max", 1.0)
    c.x3min, c.x3max = kw.get("x3min", 0.0), kw.get("x3max", 1.0)
    c.gammad = kw.get("gammad", 1.4)
    rv = kw.get("recon_vertical", ("weno5", False, False))
    rh = kw.get("recon_horizontal", rv)
    c.recon_type[0], c.recon_scale[0], c.recon_shock[0] = INTERP[rv[0]], int(rv[1]), int(rv[2])
    c.recon_type[1], c.recon_scale[1], c.recon_shock[1] = INTERP[rh[0]], int(rh[1]), int(rh[2])
    rs = kw.get("riemann", "lmars")
    if isinstance(rs, str):
        rs = (rs, rs, rs)
    for d in range(3):
        c.riemann[d] = RIEMANN[rs[d]]
    c.limiter = int(kw.get("limiter", False))
    c.density_floor = kw.get("density_floor", 1e-10)
    c.pressure_floor = kw.get("pressure_floor", 1e-10)
    c.temperature_floor = kw.get("temperature_floor", 20.0)
    c.weight = kw.get("weight", 29e-3)
    c.grav1 = kw.get("grav1", 0.0)
    c.grav2 = kw.get("grav2", 0.0)
    c.grav3 = kw.get("grav3", 0.0)
    c.non_hydrostatic = kw.get("non_hydrostatic", 1.0)
    c.implicit_scheme = kw.get("implicit_scheme", 0)
    df = kw.get("disable_flux", (0, 0, 0))
    for d in range(3):
        c.disable_flux[d] = int(df[d])
    bcs = kw.get("bc", ("reflecting",) * 6)
    active = [c.nx1 > 1, c.nx1 > 1, c.nx2 > 1, c.nx2 > 1, c.nx3 > 1, c.nx3 > 1]
    for d in range(6):
        c.bc[d] = BC[bcs[d]] if active[d] else -1
    c.coord = COORD[kw.get("coord", "cartesian")]
    return c


def shape(c: OcConfig):
    g = c.nghost
    n1 = c.nx1 + 2 * g if c.nx1 > 1 else 1
    n2 = c.nx2 + 2 * g if c.nx2 > 1 else 1
    n3 = c.nx3 + 2 * g if c.nx3 > 1 else 1
    return n3, n2, n1


def interior(c: OcConfig):
    g = c.nghost
    s1 = slice(g, -g) if c.nx1 > 1 else slice(None)
    s2 = slice(g, -g) if c.nx2 > 1 else slice(None)
    s3 = slice(g, -g) if c.nx3 > 1 else slice(None)
    return (slice(None), s3, s2, s1)


def cartesian_widths(c: OcConfig):
    n3, n2, n1 = shape(c)
    d1 = np.full(n1, (c.x1max - c.x1min) / c.nx1)
    d2 = np.full(n2, (c.x2max - c.x2min) / c.nx2)
    d3 = np.full(n3, (c.x3max - c.x3min) / c.nx3)
    return d1, d2, d3


# ---- thin array-level wrappers ----

def interp_weno5(*a):
    return lib().oc_interp_weno5(*map(float, a))


def interp_weno3(*a):
    return lib().oc_interp_weno3(*map(float, a))


def interp_cp5(*a):
    return lib().oc_interp_cp5(*map(float, a))


def interp_cp3(*a):
    return lib().oc_interp_cp3(*map(float, a))


def interp_plm(*a):
    return lib().oc_interp_plm(*map(float, a))


def interp_window(type, right, phi, scale=False):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    return lib().oc_interp_window(INTERP[type], int(right), int(scale), _p(phi))


def cons2prim(c, u):
    _check_field(c, u, 5)
    u = np.ascontiguousarray(u, dtype=np.float64).copy()
    w = np.empty_like(u)
    lib().oc_cons2prim(C.byref(c), _p(u), _p(w), u[0].size)
    return w


def prim2cons(c, w):
    _check_field(c, w, 5)
    w = np.ascontiguousarray(w, dtype=np.float64).copy()
    u = np.empty_like(w)
    lib().oc_prim2cons(C.byref(c), _p(w), _p(u), w[0].size)
    return u


def reconstruct(w, dim, type="weno5", scale=False, shock=True, has_parent=True,
                limiter=False, dfloor=1e-10, pfloor=1e-10):
    w = np.ascontiguousarray(w, dtype=np.float64)
    nvar, n3, n2, n1 = w.shape
    wlr = np.empty((2,) + w.shape)
    lib().oc_reconstruct(_p(w), _p(wlr), nvar, n3, n2, n1, dim, INTERP[type],
                         int(scale), int(shock), int(has_parent), int(limiter),
                         dfloor, pfloor)
    return wlr


def riemann(solver, wl, wr, dim, gammad=1.4):
    wl = np.ascontiguousarray(wl, dtype=np.float64).copy()
    wr = np.ascontiguousarray(wr, dtype=np.float64).copy()
    nvar, n3, n2, n1 = wl.shape
    flx = np.zeros_like(wl)
    fp = np.zeros(wl.shape[1:])
    lib().oc_riemann(RIEMANN[solver], _p(wl), _p(wr), _p(flx), _p(fp), nvar, n3, n2,
                     n1, dim, gammad)
    return flx, fp


def divergence(c, f1, f2, f3):
    for f in (f1, f2, f3):
        if f is not None:
            _check_field(c, f, 5)
    n3, n2, n1 = shape(c)
    d1, d2, d3 = cartesian_widths(c)
    div = np.zeros((5, n3, n2, n1))
    lib().oc_divergence(_p(f1), _p(f2), _p(f3), _p(div), 5, n3, n2, n1, c.nghost,
                        _p(d1), _p(d2), _p(d3))
    return div


def coord_forward(c, w, f1, f2, f3, fp1=None):
    """CoordinateImpl::forward: divergence plus, on spherical-polar blocks,
    the geometric momentum sources (every cell, ghosts included)."""
    for f in (w, f1, f2, f3):
        if f is not None:
            _check_field(c, f, 5)
    n3, n2, n1 = shape(c)
    if fp1 is not None:
        fp1 = np.ascontiguousarray(fp1, dtype=np.float64)
        assert fp1.shape == (n3, n2, n1)
    div = np.zeros((5, n3, n2, n1))
    lib().oc_coord_forward(C.byref(c), _p(w), _p(f1), _p(f2), _p(f3), _p(fp1), _p(div))
    return div


_GEOM = {"x1v": 0, "x2v": 1, "dx1v": 2, "vol": 3, "area1": 4, "area2": 5, "area3": 6,
         "width1": 7, "width2": 8, "width3": 9, "cos_cell": 10, "sin_cell": 11,
         "cos_face2": 12, "sin_face2": 13, "cos_face3": 14, "sin_face3": 15,
         "x_src": 16, "y_src": 17}


def geometry(c, what):
    """Coordinate arrays of the block: x1v, x2v, dx1v (1-D), vol, area1/2/3
    (face arrays, one extra face along their axis) and the CFL widths."""
    n3, n2, n1 = shape(c)
    shp = {"x1v": (n1,), "x2v": (n2,), "dx1v": (n1,), "area1": (n3, n2, n1 + 1),
           "area2": (n3, n2 + 1, n1), "area3": (n3 + 1, n2, n1),
           "cos_cell": (n3, n2), "sin_cell": (n3, n2), "cos_face2": (n3, n2 + 1),
           "sin_face2": (n3, n2 + 1), "cos_face3": (n3 + 1, n2),
           "sin_face3": (n3 + 1, n2)}.get(what, (n3, n2, n1))
    out = np.zeros(shp)
    lib().oc_geometry(C.byref(c), _GEOM[what], _p(out))
    return out


def gno_prim2local(c, w, dim):
    """Interface states of dim's faces in the gnomonic face frame (copy)."""
    w = np.ascontiguousarray(w, dtype=np.float64).copy()
    lib().oc_gno_prim2local(C.byref(c), _p(w), dim)
    return w


def gno_flux2global(c, f, dim):
    """Face-frame fluxes of dim's faces back to the panel components (copy)."""
    f = np.ascontiguousarray(f, dtype=np.float64).copy()
    lib().oc_gno_flux2global(C.byref(c), _p(f), dim)
    return f


def hydro_forward(c, dt, u, w=None):
    """Returns (du, w, (F1, F2, F3)); u and w are copied (u is limited in place
    by the reference only when the EOS limiter is on)."""
    _check_field(c, u, 5)
    u = np.ascontiguousarray(u, dtype=np.float64).copy()
    w = np.empty_like(u) if w is None else np.ascontiguousarray(w, dtype=np.float64).copy()
    du = np.empty_like(u)
    F = [np.zeros_like(u) for _ in range(3)]
    lib().oc_hydro_forward(C.byref(c), dt, _p(u), _p(w), _p(du), _p(F[0]), _p(F[1]),
                           _p(F[2]))
    return du, w, F


def implicit_correction(c, dt, du, w):
    _check_field(c, du, 5)
    du = np.ascontiguousarray(du, dtype=np.float64).copy()
    w = np.ascontiguousarray(w, dtype=np.float64)
    lib().oc_implicit_correction(C.byref(c), dt, _p(du), _p(w))
    return du


def vic_column_matrices(c, dt, w, k, j):
    _check_field(c, w, 5)
    N = 5 if (c.implicit_scheme >> 3) & 1 else 3
    nl = c.nx1
    a = np.zeros((nl, N, N))
    b = np.zeros((nl, N, N))
    cc = np.zeros((nl, N, N))
    w = np.ascontiguousarray(w, dtype=np.float64)
    lib().oc_vic_column_matrices(C.byref(c), dt, _p(w), k, j, _p(a), _p(b), _p(cc))
    return a, b, cc


def flux_jacobian5(w, gammad=1.4):
    """dF1/dU (5x5) at the primitive state w."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    J = np.empty((5, 5))
    lib().oc_flux_jacobian5(gammad - 1.0, _p(w), _p(J))
    return J


def roe_state(wl, wr, gammad=1.4):
    """Roe-averaged primitive state and sound speed of one interface."""
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    prim = np.empty(5)
    cs = C.c_double()
    lib().oc_roe_state(gammad, _p(wl), _p(wr), _p(prim), C.byref(cs))
    return prim, cs.value


def roe_eigensystem(wl, wr, gammad=1.4):
    """(R, R^-1, signed eigenvalues, |A|) of the interface between wl and wr."""
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    R, Ri, A = np.empty((5, 5)), np.empty((5, 5)), np.empty((5, 5))
    lam = np.empty(5)
    lib().oc_roe_eigensystem(gammad, _p(wl), _p(wr), _p(R), _p(Ri), _p(lam), _p(A))
    return R, Ri, lam, A


def vic_assemble(N, wcol, area, vol, dt, grav, gammad=1.4):
    """Blocks a, b, c ([nl, N, N]) of one column; wcol is [5, nl + 2] with one
    ghost cell at each end, area has nl + 1 faces and vol nl cells."""
    wcol = np.ascontiguousarray(wcol, dtype=np.float64)
    nl = wcol.shape[1] - 2
    area = np.ascontiguousarray(area, dtype=np.float64)
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    assert area.shape == (nl + 1,) and vol.shape == (nl,)
    a, b, cc = (np.zeros((nl, N, N)) for _ in range(3))
    first = C.cast(wcol.ctypes.data + 8, C.POINTER(C.c_double))
    lib().oc_vic_assemble(N, first, nl + 2, nl, gammad, _p(area), _p(vol), dt, grav,
                          _p(a), _p(b), _p(cc))
    return a, b, cc


PLANTS = {"c_jac_i": 1, "no_bottom_fold": 2, "phi_place": 3, "b_abs_plus": 4,
          "rinv_sign": 5, "fluxjac": 6, "roe_velocity": 7, "roe_upwind_swap": 8,
          "roe_llf_speed": 9, "gno_flux_cos": 10, "gno_xov_sin": 11, "gno_no_raise": 12,
          "gno_ang3": 13}


class plant:
    """Context manager that switches on one planted mistake in the oracle
    (tests only: shows that a check fails on a plausible error)."""

    def __init__(self, name):
        self.code = PLANTS[name]

    def __enter__(self):
        C.c_int.in_dll(lib(), "oc_plant").value = self.code
        return self

    def __exit__(self, *exc):
        C.c_int.in_dll(lib(), "oc_plant").value = 0
        return False


def inverse3(m):
    m = np.ascontiguousarray(m, dtype=np.float64)
    out = np.empty((3, 3))
    lib().oc_inverse3(_p(m), _p(out))
    return out


def apply_bc(c, var, var_type=CONSERVED):
    _check_field(c, var, None)
    var = np.ascontiguousarray(var, dtype=np.float64).copy()
    lib().oc_apply_bc(C.byref(c), _p(var), var.shape[0], var_type)
    return var


def max_time_step(c, w, cfl):
    _check_field(c, w, 5)
    w = np.ascontiguousarray(w, dtype=np.float64)
    return lib().oc_max_time_step(C.byref(c), _p(w), cfl)


def initialize(c, w):
    _check_field(c, w, 5)
    w = np.ascontiguousarray(w, dtype=np.float64).copy()
    u = np.empty_like(w)
    lib().oc_initialize(C.byref(c), _p(w), _p(u))
    return w, u


def rk3_step(c, dt, u, w=None):
    _check_field(c, u, 5)
    u = np.ascontiguousarray(u, dtype=np.float64).copy()
    w = np.empty_like(u) if w is None else np.ascontiguousarray(w, dtype=np.float64).copy()
    lib().oc_rk3_step(C.byref(c), dt, _p(u), _p(w))
    return u, w


def set_threads(n: int):
    lib().oc_set_threads(int(n))


# ---- the reference's problem set-ups (synthetic, deterministic) ----

def sod_config(nx=256):
    """1-D Sod shock tube: WENO5 (shock: true) + HLLC, outflow boundaries."""
    return make_config(nx1=nx, nghost=3, x1min=0.0, x1max=1.0, gammad=1.4,
                       recon_vertical=("weno5", False, True), riemann="hllc",
                       bc=("outflow",) * 6)


def sod_init(c):
    n3, n2, n1 = shape(c)
    x = c.x1min + (np.arange(n1) - c.nghost + 0.5) * (c.x1max - c.x1min) / c.nx1
    w = np.zeros((5, n3, n2, n1))
    left = x < 0.5
    w[IDN] = np.where(left, 1.0, 0.125)
    w[IPR] = np.where(left, 1.0, 0.1)
    return w


def straka_config(nx2=512, nx1=64, implicit_scheme=1, riemann="lmars"):
    """2-D Straka density current: x1 vertical [0, 6.4 km], x2 periodic
    [0, 25.6 km]; cp = 1004, R = 287; WENO5 (shock: false) + LMARS; gravity
    9.81; reflecting x1."""
    R, cp = 287.0, 1004.0
    return make_config(nx1=nx1, nx2=nx2, nghost=3, x1min=0.0, x1max=6.4e3,
                       x2min=0.0, x2max=25.6e3, x3min=0.0, x3max=25.6e3,
                       gammad=cp / (cp - R), weight=8.31446261815324 / R,
                       recon_vertical=("weno5", False, False), riemann=riemann,
                       grav1=-9.81, implicit_scheme=implicit_scheme,
                       bc=("reflecting", "reflecting", "periodic", "periodic",
                           "periodic", "periodic"))


def straka_init(c, R=287.0, cp=1004.0, p0=1e5, Ts=300.0, dT=-15.0, xc=0.0,
                xr=4.0e3, zc=3.0e3, zr=2.0e3, grav=9.81):
    """examples/straka.cpp:53-106 initial state (primitive)."""
    n3, n2, n1 = shape(c)
    dx1 = (c.x1max - c.x1min) / c.nx1
    dx2 = (c.x2max - c.x2min) / c.nx2
    x1v = c.x1min + (np.arange(n1) - c.nghost + 0.5) * dx1
    x2v = c.x2min + (np.arange(n2) - c.nghost + 0.5) * dx2
    X2, X1 = np.meshgrid(x2v, x1v, indexing="ij")
    L = np.sqrt(((X2 - xc) / xr) ** 2 + ((X1 - zc) / zr) ** 2)
    temp = Ts - grav * X1 / cp
    w = np.zeros((5, n3, n2, n1))
    w[IPR, 0] = p0 * np.power(temp / Ts, cp / R)
    temp = temp + np.where(L <= 1, dT * (np.cos(L * np.pi) + 1.0) / 2.0, 0.0)
    w[IDN, 0] = w[IPR, 0] / (R * temp)
    return w
