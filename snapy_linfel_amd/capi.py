"""ctypes binding of the C ABI in include/snapx.h (libsnapx.so).

This is the Python-side FFI a caller binds to the drop-in library; arrays are
torch device tensors (the data carrier), passed as raw pointers together with
torch's current HIP stream.  Every call checks the returned status and raises
``SnapxError`` with ``snapx_last_error()`` on failure.  There is no CPU
fallback: without the built library or without a GPU the calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsnapx.so")

F32, F64 = 0, 1
INTERP = {"dc": 0, "plm": 1, "cp3": 2, "cp5": 3, "weno3": 4, "weno5": 5}
RIEMANN = {"roe": 0, "lmars": 1, "hllc": 2}
BC = {None: -1, "none": -1, "reflecting": 0, "outflow": 1, "periodic": 2, "custom": 3}
COORD = {"cartesian": 0, "spherical-polar": 1, "gnomonic-equiangle": 2}
PRIMITIVE, CONSERVED, SCALAR = 0, 1, 2


class SnapxError(RuntimeError):
    pass


class Config(C.Structure):
    """snapx_config (include/snapx.h)."""
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

    @property
    def shape(self):
        g = self.nghost
        n1 = self.nx1 + 2 * g if self.nx1 > 1 else 1
        n2 = self.nx2 + 2 * g if self.nx2 > 1 else 1
        n3 = self.nx3 + 2 * g if self.nx3 > 1 else 1
        return (5, n3, n2, n1)

    @property
    def n_interior(self):
        return self.nx1 * self.nx2 * self.nx3


def make_config(**kw) -> Config:
    """Config from keyword options; names follow the hot-path YAML keys."""
    c = Config()
    c.nx1, c.nx2, c.nx3 = kw.get("nx1", 1), kw.get("nx2", 1), kw.get("nx3", 1)
    c.nghost = kw.get("nghost", 3)
    c.x1min, c.x1max = kw.get("x1min", 0.0), kw.get("x1max", 1.0)
    c.x2min, c.x2max = kw.get("x2min", 0.0), kw.get("x2max", 1.0)
    c.x3min, c.x3max = kw.get("x3min", 0.0), kw.get("x3max", 1.0)
    c.gammad = kw.get("gammad", 1.4)
    rv = kw.get("recon_vertical", ("weno5", False, False))
    rh = kw.get("recon_horizontal", rv)
    for r, (t, sc, sh) in enumerate((rv, rh)):
        c.recon_type[r], c.recon_scale[r], c.recon_shock[r] = INTERP[t], int(sc), int(sh)
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
    c.grav1, c.grav2, c.grav3 = kw.get("grav1", 0.0), kw.get("grav2", 0.0), kw.get("grav3", 0.0)
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


_lib = None

# name -> (restype, argtypes)
_P, _I, _D, _Z, _L = C.c_void_p, C.c_int, C.c_double, C.c_size_t, C.c_int64
_CFG = C.POINTER(Config)
SIGNATURES = {
    "snapx_last_error": (C.c_char_p, []),
    "snapx_version": (_I, []),
    "snapx_device_count": (_I, []),
    "snapx_ideal_gas_cons2prim": (_I, [_CFG, _P, _P, _I, _P]),
    "snapx_ideal_gas_prim2cons": (_I, [_CFG, _P, _P, _I, _P]),
    "snapx_reconstruct": (_I, [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _D, _D, _I, _I,
                               _I, _P]),
    "snapx_riemann": (_I, [_I, _P, _P, _P, _P, _I, _I, _I, _I, _I, _D, _I, _P]),
    "snapx_divergence": (_I, [_CFG, _P, _P, _P, _P, _I, _P]),
    "snapx_coord_forward": (_I, [_CFG, _P, _P, _P, _P, _P, _P, _I, _P]),
    "snapx_revise_x1_ghost": (_I, [_CFG, _P, _I, _P]),
    "snapx_apply_bc": (_I, [_CFG, _P, _I, _I, _I, _P]),
    "snapx_rk_combine": (_I, [_P, _P, _P, _P, _D, _D, _D, _L, _I, _P]),
    "snapx_implicit_workspace_bytes": (_Z, [_CFG]),
    "snapx_implicit_correction": (_I, [_CFG, _D, _P, _P, _P, _Z, _I, _P]),
    "snapx_hydro_workspace_bytes": (_Z, [_CFG, _I]),
    "snapx_hydro_forward": (_I, [_CFG, _D, _P, _P, _P, _P, _P, _P, _P, _Z, _I, _P]),
    "snapx_rk3_workspace_bytes": (_Z, [_CFG, _I]),
    "snapx_rk3_step": (_I, [_CFG, _D, _P, _P, _P, _Z, _I, _P]),
    "snapx_rk3_step_dt": (_I, [_CFG, _P, _P, _P, _P, _Z, _I, _P]),
    "snapx_max_time_step": (_I, [_CFG, _P, _D, _P, _I, _P]),
    "snapx_check_state": (_I, [_CFG, _P, _P, _I, _P]),
    "snapx_fused_workspace_bytes": (_Z, [_CFG]),
    "snapx_fused_rk3_step": (_I, [_CFG, _D, _P, _P, _Z, _P]),
    "snapx_fused_rk3_step_dt": (_I, [_CFG, _P, _P, _P, _Z, _P]),
    "snapx_fused_rk3_steps": (_I, [_CFG, _D, _I, _P, _P, _Z, _P, _P]),
    "snapx_fused_rk3_steps_static": (_I, [_CFG, _D, _I, _P, _P]),
    "snapx_fused_background_bytes": (_Z, [_CFG]),
    "snapx_fused_background": (_I, [_CFG, _P, _P, _P, _Z, _P]),
    "snapx_fused_split": (_I, [_CFG, _P, _P, _I, _P, _P]),
    "snapx_fused_join": (_I, [_CFG, _P, _P, _P, _I, _P]),
    "snapx_fused_rk3_steps_bg": (_I, [_CFG, _P, _D, _I, _P, _P, _Z, _P, _P]),
    "snapx_adjoint_workspace_bytes": (_Z, [_CFG]),
    "snapx_adjoint_rk3_step": (_I, [_CFG, _D, _P, _P, _P, _Z, _P]),
    "snapx_adjoint_tape_bytes": (_Z, [_CFG]),
    "snapx_adjoint_work_bytes": (_Z, [_CFG]),
    "snapx_adjoint_record": (_I, [_CFG, _D, _P, _P, _P, _Z, _P, _Z, _P]),
    "snapx_adjoint_sweep": (_I, [_CFG, _D, _P, _Z, _P, _P, _Z, _P]),
    "snapx_adjoint_f32_workspace_bytes": (_Z, [_CFG]),
    "snapx_adjoint_f32_rk3_step": (_I, [_CFG, _D, _P, _P, _P, _Z, _P]),
    "snapx_adjoint_f32_tape_bytes": (_Z, [_CFG]),
    "snapx_adjoint_f32_work_bytes": (_Z, [_CFG]),
    "snapx_adjoint_f32_record": (_I, [_CFG, _D, _P, _P, _P, _Z, _P, _Z, _P]),
    "snapx_adjoint_f32_sweep": (_I, [_CFG, _D, _P, _Z, _P, _P, _Z, _P]),
    "snapx_batch_workspace_bytes": (_Z, [_CFG, _I]),
    "snapx_batch_plan": (_I, [_CFG, _I, _D, _I, _I, _I, _P, _Z, _P]),
    "snapx_batch_rk3_steps": (_I, [_CFG, _P, _I, _P, _P, _Z, _P, _P]),
    "snapx_batch_autotune": (_I, [_CFG, _I, _D, _P, _P, _Z, _I, _P, _P, _P]),
    "snapx_batch_set_dt": (_I, [_P, _P, _P, _Z, _P]),
    "snapx_rk3_attempts_workspace_bytes": (_Z, [_CFG, _I]),
    "snapx_rk3_attempts": (_I, [_CFG, _P, _I, _I, _P, _P, _P, _P, _Z, _I, _P]),
    "snapx_fused_attempts_workspace_bytes": (_Z, [_CFG]),
    "snapx_fused_attempts": (_I, [_CFG, _P, _I, _P, _P, _P, _Z, _P]),
    "snapx_fused_graph_create": (_I, [_CFG, _D, _I, _P, _P, _Z, _P]),
    "snapx_batch_graph_create": (_I, [_CFG, _P, _I, _P, _P, _Z, _P]),
    "snapx_graph_launch": (_I, [_P, _P]),
    "snapx_graph_destroy": (_I, [_P]),
}


class BatchInfo(C.Structure):
    """snapx_batch_info (include/snapx.h)."""
    _fields_ = [("nmembers", C.c_int), ("tile_x1", C.c_int), ("tile_x2", C.c_int),
                ("tile_x3", C.c_int), ("total_tiles", C.c_int), ("total_cols", C.c_int),
                ("vic_size", C.c_int), ("total_elems", C.c_int64),
                ("scratch_doubles", C.c_int64), ("rhs_smem", C.c_size_t), ("dt", C.c_double),
                ("col_groups", C.c_int), ("col_groups2", C.c_int)]


class Timing(C.Structure):
    """snapx_timing (include/snapx.h)."""
    _fields_ = [("rhs_ms", C.c_double), ("column_ms", C.c_double), ("bc_ms", C.c_double),
                ("rhs_launches", C.c_int64), ("column_launches", C.c_int64),
                ("bc_launches", C.c_int64)]


def lib():
    """Load libsnapx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SnapxError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int, what: str):
    if status != 0:
        msg = lib().snapx_last_error().decode()
        raise SnapxError(f"{what} failed ({status}): {msg}")


# ---- torch helpers ----

def _torch():
    import torch
    return torch


def stream_ptr():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    if t is None:
        return None
    torch = _torch()
    assert t.is_cuda, "snapx expects device tensors"
    assert t.is_contiguous(), "snapx expects contiguous tensors"
    return C.c_void_p(t.data_ptr())


def dtype_code(t):
    torch = _torch()
    if t.dtype == torch.float64:
        return F64
    if t.dtype == torch.float32:
        return F32
    raise SnapxError(f"unsupported dtype {t.dtype}")


def workspace(nbytes, device="cuda"):
    torch = _torch()
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def cons2prim(cfg, u, w):
    check(lib().snapx_ideal_gas_cons2prim(C.byref(cfg), ptr(u), ptr(w), dtype_code(u),
                                          stream_ptr()), "cons2prim")


def prim2cons(cfg, w, u):
    check(lib().snapx_ideal_gas_prim2cons(C.byref(cfg), ptr(w), ptr(u), dtype_code(w),
                                          stream_ptr()), "prim2cons")


def reconstruct(w, wlr, dim, type="weno5", scale=False, shock=True, has_parent=True,
                limiter=False, dfloor=1e-10, pfloor=1e-10, revise_x1_lr=False, nghost=3):
    nvar, n3, n2, n1 = w.shape
    check(lib().snapx_reconstruct(ptr(w), ptr(wlr), nvar, n3, n2, n1, dim, INTERP[type],
                                  int(scale), int(shock), int(has_parent), int(limiter), dfloor,
                                  pfloor, int(revise_x1_lr), nghost, dtype_code(w),
                                  stream_ptr()), "reconstruct")


def riemann(solver, wl, wr, flx, fp, dim, gammad=1.4):
    nvar, n3, n2, n1 = wl.shape
    check(lib().snapx_riemann(RIEMANN[solver], ptr(wl), ptr(wr), ptr(flx), ptr(fp), nvar, n3,
                              n2, n1, dim, gammad, dtype_code(wl), stream_ptr()), "riemann")


def divergence(cfg, f1, f2, f3, div):
    check(lib().snapx_divergence(C.byref(cfg), ptr(f1), ptr(f2), ptr(f3), ptr(div),
                                 dtype_code(div), stream_ptr()), "divergence")


def coord_forward(cfg, w, f1, f2, f3, fp1, div):
    """CoordinateImpl::forward: divergence plus the spherical-polar sources."""
    check(lib().snapx_coord_forward(C.byref(cfg), ptr(w), ptr(f1), ptr(f2), ptr(f3), ptr(fp1),
                                    ptr(div), dtype_code(div), stream_ptr()), "coord_forward")


def revise_x1_ghost(cfg, w):
    check(lib().snapx_revise_x1_ghost(C.byref(cfg), ptr(w), dtype_code(w), stream_ptr()),
          "revise_x1_ghost")


def apply_bc(cfg, var, kind=CONSERVED):
    check(lib().snapx_apply_bc(C.byref(cfg), ptr(var), var.shape[0], kind, dtype_code(var),
                               stream_ptr()), "apply_bc")


def rk_combine(u, u0, u1, du, w0, w1, w2):
    check(lib().snapx_rk_combine(ptr(u), ptr(u0), ptr(u1), ptr(du), w0, w1, w2, u.numel(),
                                 dtype_code(u), stream_ptr()), "rk_combine")


def implicit_correction(cfg, dt, du, w, ws=None):
    n = lib().snapx_implicit_workspace_bytes(C.byref(cfg))
    ws = workspace(n, du.device) if ws is None else ws
    check(lib().snapx_implicit_correction(C.byref(cfg), dt, ptr(du), ptr(w), ptr(ws), n,
                                          dtype_code(du), stream_ptr()), "implicit_correction")


def hydro_forward(cfg, dt, u, w, du, f1=None, f2=None, f3=None, ws=None):
    n = lib().snapx_hydro_workspace_bytes(C.byref(cfg), dtype_code(u))
    ws = workspace(n, u.device) if ws is None else ws
    check(lib().snapx_hydro_forward(C.byref(cfg), dt, ptr(u), ptr(w), ptr(du), ptr(f1), ptr(f2),
                                    ptr(f3), ptr(ws), ws.numel(), dtype_code(u), stream_ptr()),
          "hydro_forward")


def rk3_workspace(cfg, dtype, device="cuda"):
    return workspace(lib().snapx_rk3_workspace_bytes(C.byref(cfg), dtype), device)


def rk3_step(cfg, dt, u, w, ws=None):
    ws = rk3_workspace(cfg, dtype_code(u), u.device) if ws is None else ws
    check(lib().snapx_rk3_step(C.byref(cfg), dt, ptr(u), ptr(w), ptr(ws), ws.numel(),
                               dtype_code(u), stream_ptr()), "rk3_step")


def rk3_step_dt(cfg, dt, u, w, ws=None):
    """rk3_step with the time step read from the one-element float64 device
    tensor ``dt``."""
    _check_dt(dt, 1, u.device)
    ws = rk3_workspace(cfg, dtype_code(u), u.device) if ws is None else ws
    check(lib().snapx_rk3_step_dt(C.byref(cfg), ptr(dt), ptr(u), ptr(w), ptr(ws), ws.numel(),
                                  dtype_code(u), stream_ptr()), "rk3_step_dt")


def _check_dt(dt, n, device):
    torch = _torch()
    if dt.dtype != torch.float64 or dt.numel() < n or dt.device != device:
        raise SnapxError(f"dt must be a float64 device tensor of at least {n} elements")


def max_time_step(cfg, w, cfl, out):
    check(lib().snapx_max_time_step(C.byref(cfg), ptr(w), cfl, ptr(out), dtype_code(w),
                                    stream_ptr()), "max_time_step")


REDO_DENSITY, REDO_PRESSURE, STATE_NONFINITE = 1, 2, 4


def check_state(cfg, u, flag):
    """Writes the positivity verdict of ``u`` (REDO_* / STATE_NONFINITE bits)
    into the one-element int32 device tensor ``flag``; no synchronisation."""
    torch = _torch()
    if flag.dtype != torch.int32 or flag.numel() < 1 or flag.device != u.device:
        raise ValueError("check_state: flag must be an int32 tensor on u's device")
    check(lib().snapx_check_state(C.byref(cfg), ptr(u), ptr(flag), dtype_code(u), stream_ptr()),
          "check_state")


def _adj(name, dtype):
    """The FP64 (float64) or FP32-storage (float32) adjoint entry point."""
    torch = _torch()
    if dtype == torch.float64:
        return getattr(lib(), "snapx_adjoint_" + name)
    if dtype == torch.float32:
        return getattr(lib(), "snapx_adjoint_f32_" + name)
    raise SnapxError(f"adjoint: state dtype {dtype}, expected float64 or float32")


def _adj_pair(cfg, a, b, what):
    torch = _torch()
    _check_state(cfg, a, (torch.float64, torch.float32))
    if b is not None:
        _check_state(cfg, b, (a.dtype,))
        if b.device != a.device:
            raise SnapxError(f"{what}: tensors on different devices")


def adjoint_workspace(cfg, device="cuda", dtype=None):
    dtype = _torch().float64 if dtype is None else dtype
    return workspace(_adj("workspace_bytes", dtype)(C.byref(cfg)), device)


def adjoint_rk3_step(cfg, dt, u_in, lam, ws=None):
    """lam <- (d u_out / d u_in)^T lam for one plain-path RK3 step (FP64, or
    FP32 storage with FP64 column solves when the tensors are float32)."""
    _adj_pair(cfg, u_in, lam, "adjoint_rk3_step")
    ws = adjoint_workspace(cfg, u_in.device, u_in.dtype) if ws is None else ws
    check(_adj("rk3_step", u_in.dtype)(C.byref(cfg), dt, ptr(u_in), ptr(lam), ptr(ws),
                                       ws.numel(), stream_ptr()), "adjoint_rk3_step")


def adjoint_tape_bytes(cfg, dtype=None):
    dtype = _torch().float64 if dtype is None else dtype
    return _adj("tape_bytes", dtype)(C.byref(cfg))


def adjoint_work(cfg, device="cuda", dtype=None):
    dtype = _torch().float64 if dtype is None else dtype
    return workspace(_adj("work_bytes", dtype)(C.byref(cfg)), device)


def adjoint_record(cfg, dt, u_in, u_out, tape, work):
    """Forward recompute of one step from u_in into ``tape`` (a uint8
    device tensor of adjoint_tape_bytes for u_in's dtype); u_out (or None)
    gets the step's output."""
    _adj_pair(cfg, u_in, u_out, "adjoint_record")
    check(_adj("record", u_in.dtype)(C.byref(cfg), dt, ptr(u_in), ptr(u_out), ptr(tape),
                                     tape.numel(), ptr(work), work.numel(), stream_ptr()),
          "adjoint_record")


def adjoint_sweep(cfg, dt, tape, lam, work):
    """lam <- (d u_out / d u_in)^T lam from a recorded tape (of lam's
    dtype)."""
    _adj_pair(cfg, lam, None, "adjoint_sweep")
    check(_adj("sweep", lam.dtype)(C.byref(cfg), dt, ptr(tape), tape.numel(), ptr(lam),
                                   ptr(work), work.numel(), stream_ptr()), "adjoint_sweep")


def fused_workspace(cfg, device="cuda"):
    return workspace(lib().snapx_fused_workspace_bytes(C.byref(cfg)), device)


def fused_rk3_steps(cfg, dt, nsteps, u, ws=None, timing=None):
    _check_state(cfg, u, (_torch().float32,))
    ws = fused_workspace(cfg, u.device) if ws is None else ws
    tp = C.byref(timing) if timing is not None else None
    check(lib().snapx_fused_rk3_steps(C.byref(cfg), dt, nsteps, ptr(u), ptr(ws), ws.numel(), tp,
                                      stream_ptr()), "fused_rk3_steps")


class Background:
    """FP64 hydrostatic background of the FP32 perturbation path: rho_bar and
    p_bar per x1 cell level (ghosts included, length nc1), horizontally
    uniform and at rest.  ``split`` / ``join`` convert a total state to the
    FP32 perturbation and back; ``steps`` advances the perturbation."""

    def __init__(self, cfg, rho_bar, p_bar, device="cuda"):
        import numpy as np
        n1 = cfg.shape[3]
        rb = np.ascontiguousarray(rho_bar, dtype=np.float64)
        pb = np.ascontiguousarray(p_bar, dtype=np.float64)
        if rb.shape != (n1,) or pb.shape != (n1,):
            raise SnapxError(f"background: rho_bar and p_bar need {n1} levels")
        nb = lib().snapx_fused_background_bytes(C.byref(cfg))
        if nb == 0:
            raise SnapxError("background: " + lib().snapx_last_error().decode())
        self.cfg = cfg
        self.buf = workspace(nb, device)
        check(lib().snapx_fused_background(C.byref(cfg), rb.ctypes.data_as(C.c_void_p),
                                           pb.ctypes.data_as(C.c_void_p), ptr(self.buf), nb,
                                           stream_ptr()), "fused_background")

    def split(self, u, out=None):
        """FP32 perturbation u - ubar of the total state u (float or double)."""
        torch = _torch()
        _check_state(self.cfg, u, (torch.float32, torch.float64))
        out = torch.empty(u.shape, dtype=torch.float32, device=u.device) if out is None else out
        check(lib().snapx_fused_split(C.byref(self.cfg), ptr(self.buf), ptr(u), dtype_code(u),
                                      ptr(out), stream_ptr()), "fused_split")
        return out

    def join(self, up, out):
        """out = ubar + up with the external boundaries applied (out float or
        double)."""
        torch = _torch()
        _check_state(self.cfg, up, (torch.float32,))
        _check_state(self.cfg, out, (torch.float32, torch.float64))
        check(lib().snapx_fused_join(C.byref(self.cfg), ptr(self.buf), ptr(up), ptr(out),
                                     dtype_code(out), stream_ptr()), "fused_join")
        return out

    def steps(self, dt, nsteps, up, ws=None, timing=None):
        """nsteps fused steps of the FP32 perturbation up."""
        torch = _torch()
        _check_state(self.cfg, up, (torch.float32,))
        ws = fused_workspace(self.cfg, up.device) if ws is None else ws
        tp = C.byref(timing) if timing is not None else None
        check(lib().snapx_fused_rk3_steps_bg(C.byref(self.cfg), ptr(self.buf), dt, nsteps,
                                             ptr(up), ptr(ws), ws.numel(), tp, stream_ptr()),
              "fused_rk3_steps_bg")


def _check_state(cfg, u, dtypes):
    """Rejects a tensor the kernels would misread: wrong dtype, element count
    or device."""
    n = 1
    for d in cfg.shape:
        n *= d
    if u.dtype not in dtypes:
        raise SnapxError(f"state dtype {u.dtype}, expected one of {dtypes}")
    if u.numel() != n:
        raise SnapxError(f"state has {u.numel()} elements, the block needs {n}")
    if not u.is_cuda or not u.is_contiguous():
        raise SnapxError("state must be a contiguous device tensor")


STATIC_SCRATCH_BYTES = 64 << 20
ESPACE = -3


def fused_rk3_steps_static(cfg, dt, nsteps, u):
    """fused_rk3_steps with the library's static per-GPU workspace: no
    allocation at all (calls that overlap on one GPU race)."""
    check(lib().snapx_fused_rk3_steps_static(C.byref(cfg), dt, nsteps, ptr(u), stream_ptr()),
          "fused_rk3_steps_static")


def fused_rk3_step(cfg, dt, u, ws=None):
    _check_state(cfg, u, (_torch().float32,))
    ws = fused_workspace(cfg, u.device) if ws is None else ws
    check(lib().snapx_fused_rk3_step(C.byref(cfg), dt, ptr(u), ptr(ws), ws.numel(),
                                     stream_ptr()), "fused_rk3_step")


def fused_rk3_step_dt(cfg, dt, u, ws=None):
    """fused_rk3_step with the time step read from the one-element float64
    device tensor ``dt``."""
    _check_state(cfg, u, (_torch().float32,))
    _check_dt(dt, 1, u.device)
    ws = fused_workspace(cfg, u.device) if ws is None else ws
    check(lib().snapx_fused_rk3_step_dt(C.byref(cfg), ptr(dt), ptr(u), ptr(ws), ws.numel(),
                                        stream_ptr()), "fused_rk3_step_dt")


DT_FROM_W, DT_FROM_U = 0, 1


class StepControl(C.Structure):
    """snapx_step_control (include/snapx.h)."""
    _fields_ = [("cfl", C.c_double), ("tlim", C.c_double), ("nlim", C.c_int64),
                ("max_redo", C.c_int32), ("redo_mask", C.c_int32),
                ("time", C.c_double), ("dt", C.c_double), ("dt_raw", C.c_double),
                ("cycle", C.c_int64), ("redo", C.c_int32), ("status", C.c_int32),
                ("verdict", C.c_int32), ("active", C.c_int32), ("steps", C.c_int32),
                ("redos", C.c_int32), ("restore", C.c_int32), ("reserved2", C.c_int32)]


class StepLoop:
    """The CFL time loop with redo, its state in device memory
    (snapx_step_control): ``attempts`` / ``fused_attempts`` enqueue attempts
    without synchronising, so they can be captured in a graph; ``state``
    reads the record back (synchronising)."""

    def __init__(self, cfg, cfl, tlim=float("inf"), nlim=-1, max_redo=5,
                 redo_mask=None, time=0.0, cycle=0, device="cuda"):
        torch = _torch()
        self.cfg = cfg
        self.rec = torch.zeros(C.sizeof(StepControl), dtype=torch.uint8, device=device)
        self.ws = None
        c = StepControl()
        c.cfl, c.tlim, c.nlim, c.max_redo = cfl, tlim, nlim, max_redo
        c.redo_mask = (REDO_DENSITY | REDO_PRESSURE) if redo_mask is None else redo_mask
        c.time, c.cycle = time, cycle
        self.rec.copy_(torch.frombuffer(bytearray(bytes(c)), dtype=torch.uint8))

    def state(self) -> StepControl:
        return StepControl.from_buffer_copy(bytes(self.rec.cpu().numpy().tobytes()))

    def prepare(self, u, fused=False):
        """Allocates the attempts' workspace for states like u (before a
        graph capture, so that the capture allocates nothing)."""
        if fused:
            nb = lib().snapx_fused_attempts_workspace_bytes(C.byref(self.cfg))
            if nb == 0:
                raise SnapxError("fused_attempts: " + lib().snapx_last_error().decode())
        else:
            nb = lib().snapx_rk3_attempts_workspace_bytes(C.byref(self.cfg), dtype_code(u))
        if self.ws is None or self.ws.numel() < nb:
            self.ws = workspace(nb, u.device)

    def _log(self, dt_log, n, device):
        if dt_log is not None:
            _check_dt(dt_log, n, device)
        return ptr(dt_log)

    def attempts(self, u, w, n, dt_source=DT_FROM_W, dt_log=None):
        """n attempts on the plain path (u, w float64 or float32)."""
        torch = _torch()
        _check_state(self.cfg, u, (torch.float64, torch.float32))
        _check_state(self.cfg, w, (u.dtype,))
        self.prepare(u)
        check(lib().snapx_rk3_attempts(C.byref(self.cfg), ptr(self.rec), n, dt_source, ptr(u),
                                       ptr(w), self._log(dt_log, n, u.device), ptr(self.ws),
                                       self.ws.numel(), dtype_code(u), stream_ptr()),
              "rk3_attempts")

    def fused_attempts(self, u, n, dt_log=None):
        """n attempts on the fused FP32 path."""
        _check_state(self.cfg, u, (_torch().float32,))
        self.prepare(u, fused=True)
        check(lib().snapx_fused_attempts(C.byref(self.cfg), ptr(self.rec), n, ptr(u),
                                         self._log(dt_log, n, u.device), ptr(self.ws),
                                         self.ws.numel(), stream_ptr()), "fused_attempts")


class Batch:
    """A planned batch of independent members (configs may differ in size);
    their FP32 states are concatenated in member order."""

    def __init__(self, configs, dt, tile=(0, 0, 0), device="cuda"):
        self.n = len(configs)
        self.cfgs = (Config * self.n)(*configs)
        nb = lib().snapx_batch_workspace_bytes(self.cfgs, self.n)
        if nb == 0:
            raise SnapxError("batch: " + lib().snapx_last_error().decode())
        self.ws = workspace(nb, device)
        self._graphs = weakref.WeakSet()  # live graphs captured on self.ws
        self.info = BatchInfo()
        check(lib().snapx_batch_plan(self.cfgs, self.n, dt, tile[0], tile[1], tile[2],
                                     ptr(self.ws), self.ws.numel(), C.byref(self.info)),
              "batch_plan")
        self.offsets = []
        off = 0
        for c in configs:
            self.offsets.append(off)
            n = 1
            for d in c.shape:
                n *= d
            off += n
        self.total = off

    def member(self, u, m):
        """View of member m's [5, nc3, nc2, nc1] state inside the batch array."""
        c = self.cfgs[m]
        n = 1
        for d in c.shape:
            n *= d
        return u[self.offsets[m]:self.offsets[m] + n].view(*c.shape)

    def _check(self, u):
        torch = _torch()
        if u.dtype != torch.float32 or u.numel() != self.total:
            raise SnapxError(f"batch state must be float32 with {self.total} elements "
                             f"(got {u.dtype}, {u.numel()})")
        if not u.is_cuda or not u.is_contiguous():
            raise SnapxError("batch state must be a contiguous device tensor")

    def steps(self, u, nsteps, timing=None):
        self._check(u)
        tp = C.byref(timing) if timing is not None else None
        check(lib().snapx_batch_rk3_steps(self.cfgs, C.byref(self.info), nsteps, ptr(u),
                                          ptr(self.ws), self.ws.numel(), tp, stream_ptr()),
              "batch_rk3_steps")

    def autotune(self, u, dt, trials=3):
        """Re-plan with the fastest RHS tile for these members (u unchanged);
        returns (tile, ms per RHS launch).  Refused while a graph of this
        batch exists: the graph keeps the launch shapes of the old plan."""
        self._check(u)
        if len(self._graphs):
            raise SnapxError("batch autotune: destroy the batch's graphs first (they keep "
                             "the launch shapes of the current plan)")
        ms = C.c_double(0.0)
        check(lib().snapx_batch_autotune(self.cfgs, self.n, dt, ptr(u), ptr(self.ws),
                                         self.ws.numel(), trials, stream_ptr(),
                                         C.byref(self.info), C.byref(ms)), "batch_autotune")
        return (self.info.tile_x1, self.info.tile_x2, self.info.tile_x3), ms.value

    def graph(self, u, nsteps):
        """A Graph replaying nsteps batched steps on u."""
        self._check(u)
        h = C.c_void_p()
        check(lib().snapx_batch_graph_create(self.cfgs, C.byref(self.info), nsteps, ptr(u),
                                             ptr(self.ws), self.ws.numel(), C.byref(h)),
              "batch_graph_create")
        g = Graph(h, (u, self.ws))
        self._graphs.add(g)
        return g

    def set_dt(self, dt):
        """Per-member time steps from the float64 device tensor dt
        (nmembers values), written into the member table by a kernel: the
        steps and graph replays enqueued after it use them."""
        _check_dt(dt, self.n, self.ws.device)
        check(lib().snapx_batch_set_dt(C.byref(self.info), ptr(dt), ptr(self.ws),
                                       self.ws.numel(), stream_ptr()), "batch_set_dt")


class Graph:
    """An executable HIP graph of captured steps; keeps its buffers alive."""

    def __init__(self, handle, keep):
        self.h = handle
        self._keep = keep

    def launch(self):
        check(lib().snapx_graph_launch(self.h, stream_ptr()), "graph_launch")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().snapx_graph_destroy(self.h)
            self.h = None


def fused_graph(cfg, dt, nsteps, u, ws=None):
    """A Graph replaying nsteps fused FP32 steps on u."""
    ws = fused_workspace(cfg, u.device) if ws is None else ws
    h = C.c_void_p()
    check(lib().snapx_fused_graph_create(C.byref(cfg), dt, nsteps, ptr(u), ptr(ws), ws.numel(),
                                         C.byref(h)), "fused_graph_create")
    return Graph(h, (u, ws))
