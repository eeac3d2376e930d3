"""Benchmark: cell-updates/s per SSP-RK3 step on one MI355X.

Headline workload (the largest single-GPU configuration of the metric):
3-D dry convective box, 256 x 256 x 128 cells of 100 m (x2, x3 horizontal and
periodic, 25.6 km; x1 vertical and reflecting, 12.8 km), ideal gas gamma 1.4,
R = 287, gravity 9.81, hydrostatic background of a dry adiabat from 250 K at
the ground capped by an isothermal layer at 200 K (the profile of
examples/linfel2026/uranus_dry_cart.cpp:86-101), theta += U(-0.5, 0.5) K in
the lowest 20 levels (torch CPU generator, seed 1234) at fixed pressure,
WENO5 (shock: false) with Roe in x2 / x3 and LMARS in x1, vertically
implicit 3x3 column solve (65 536 columns), SSP-RK3, dt = 0.2 s, 100 steps.
Synthetic initial state, inputs resident in HBM.

Reported:
  value        FP32 fused path carrying the state as a perturbation about the
               FP64 hydrostatic background (FP64 column solve): interior
               cells x steps / wall seconds
  fp32_total_state  the same fused step on the FP32 total state
  fp64_plain   the FP64 plain (module-by-module) GPU path on the same steps
  error        per-variable L2 / Linf relative error of FP32 fused vs FP64
               plain after the timed steps (each field against its own norm,
               and momenta also against rho_max x 300 m/s), mass drift
  oracle_check both GPU paths against the FP64 CPU oracle after 2 steps
  roofline     dominant kernel of the fused step, algorithmic bytes per
               launch / its average device time (HIP events on the stream)
  cpu_baseline the FP64 CPU oracle timed on a 2-step window of the workload
  extra        Straka 512 x 64, spherical-polar sector, 96-member ensemble,
               adjoint / forward and the adjoint sensitivity run
Run ``python bench.py`` (N = 1); the package runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
NV = 5


def straka_cfg(capi, nx2=512, nx1=64):
    R, cp = 287.0, 1004.0
    return capi.make_config(nx1=nx1, nx2=nx2, nghost=3, x1min=0.0, x1max=6.4e3, x2min=0.0,
                            x2max=25.6e3, x3min=0.0, x3max=25.6e3, gammad=cp / (cp - R),
                            weight=8.31446261815324 / R,
                            recon_vertical=("weno5", False, False), riemann="lmars",
                            grav1=-9.81, implicit_scheme=1,
                            bc=("reflecting", "reflecting", "periodic", "periodic",
                                "periodic", "periodic"))


def straka_prim(cfg, R=287.0, cp=1004.0, p0=1e5, Ts=300.0, dT=-15.0, xc=0.0, xr=4.0e3,
                zc=3.0e3, zr=2.0e3, grav=9.81):
    """Primitive initial state of examples/straka.cpp:53-106 on the block of
    cfg, ghost cells included: a neutral atmosphere (T = Ts - g z / cp,
    p = p0 (T / Ts)^(cp / R)) with a -15 K cosine bubble at (xc, zc)."""
    import numpy as np
    n = cfg.shape
    g = cfg.nghost
    dx1 = (cfg.x1max - cfg.x1min) / cfg.nx1
    dx2 = (cfg.x2max - cfg.x2min) / cfg.nx2
    z = cfg.x1min + (np.arange(n[3]) - g + 0.5) * dx1
    x = cfg.x2min + (np.arange(n[2]) - g + 0.5) * dx2
    X, Z = np.meshgrid(x, z, indexing="ij")
    L = np.sqrt(((X - xc) / xr) ** 2 + ((Z - zc) / zr) ** 2)
    T = Ts - grav * Z / cp
    w = np.zeros(n)
    w[4, 0] = p0 * np.power(T / Ts, cp / R)
    T = T + np.where(L <= 1, dT * (np.cos(L * np.pi) + 1.0) / 2.0, 0.0)
    w[0, 0] = w[4, 0] / (R * T)
    return w


def conserved_on_device(capi, cfg, w, dev):
    """W -> U and the boundary functions, on the GPU (the library's own
    initialisation path); returns the FP64 device state."""
    import torch
    wd = torch.from_numpy(w).to(dev, torch.float64)
    ud = torch.empty_like(wd)
    capi.prim2cons(cfg, wd, ud)
    capi.apply_bc(cfg, ud)
    return ud


def sector_cfg(capi, nx1=64, nx2=512):
    """Spherical-polar shell sector r in [1, 2], theta in [pi/4, 3pi/4] (one
    azimuthal cell), nondimensional: gamma 1.4, gravity -0.5, WENO5 + LMARS,
    implicit 3x3 radial solve, reflecting walls."""
    import math
    return capi.make_config(nx1=nx1, nx2=nx2, nghost=3, x1min=1.0, x1max=2.0,
                            x2min=0.25 * math.pi, x2max=0.75 * math.pi, coord="spherical-polar",
                            gammad=1.4, recon_vertical=("weno5", False, False), riemann="lmars",
                            grav1=-0.5, implicit_scheme=1,
                            bc=("reflecting", "reflecting", "reflecting", "reflecting",
                                "periodic", "periodic"))


def sector_prim(cfg):
    """Isothermal atmosphere (R T = 1, p = exp(-0.5 (r - 1))) with a warm
    Gaussian bubble (density 5 % low) at r = 1.3, theta = pi/2."""
    import numpy as np
    n = cfg.shape
    g = cfg.nghost
    dr = (cfg.x1max - cfg.x1min) / cfg.nx1
    dth = (cfg.x2max - cfg.x2min) / cfg.nx2
    r = cfg.x1min + (np.arange(n[3]) - g + 0.5) * dr
    th = cfg.x2min + (np.arange(n[2]) - g + 0.5) * dth
    TH, RR = np.meshgrid(th, r, indexing="ij")
    p = np.exp(-0.5 * (RR - 1.0))
    bub = np.exp(-(((RR - 1.3) / 0.1) ** 2 + ((TH - 0.5 * np.pi) / 0.1) ** 2))
    w = np.zeros(n)
    w[4, 0] = p
    w[0, 0] = p * (1.0 - 0.05 * bub)
    return w


def box3d_cfg(capi, n=(128, 256, 256)):
    return capi.make_config(nx1=n[0], nx2=n[1], nx3=n[2], nghost=3, x1min=0.0,
                            x1max=100.0 * n[0], x2min=0.0, x2max=100.0 * n[1], x3min=0.0,
                            x3max=100.0 * n[2], gammad=1.4, weight=8.31446261815324 / 287.0,
                            recon_vertical=("weno5", False, False),
                            riemann=("lmars", "roe", "roe"), grav1=-9.81, implicit_scheme=1,
                            bc=("reflecting", "reflecting", "periodic", "periodic", "periodic",
                                "periodic"))


def capped_background(z, T0=250.0, Tmin=200.0, R=287.0, gamma=1.4, grav=9.81, P0=1e5):
    """Hydrostatic (T, p) at heights z (torch tensor): a dry adiabat from T0 at
    z = 0, capped by an isothermal layer at Tmin above
    z_iso = cp (T0 - Tmin) / g (examples/linfel2026/uranus_dry_cart.cpp:86-101)."""
    import torch
    cp = gamma / (gamma - 1.0) * R
    t_ad = T0 - grav * z / cp
    T = torch.clamp_min(t_ad, Tmin)
    z_iso = cp * (T0 - Tmin) / grav
    p_ad = P0 * (torch.clamp_min(t_ad, Tmin) / T0) ** (cp / R)
    p_iso = P0 * (Tmin / T0) ** (cp / R) * torch.exp(-grav * (z - z_iso) / (R * Tmin))
    return T, torch.where(t_ad > Tmin, p_ad, p_iso)


def box3d_prim(cfg, dz=100.0, levels=20, seed=1234, R=287.0, P0=1e5):
    """Primitive initial state of the 3-D box (ghosts included): the capped
    background with theta += U(-0.5, 0.5) K, i.i.d. per cell, in the lowest
    `levels` interior levels, applied at fixed pressure."""
    import torch
    shp = cfg.shape
    g = cfg.nghost
    cp = cfg.gammad / (cfg.gammad - 1.0) * R
    z = (torch.arange(shp[3], dtype=torch.float64) - g + 0.5) * dz
    T, p = capped_background(z, gamma=cfg.gammad)
    gen = torch.Generator().manual_seed(seed)
    dth = torch.rand(shp[1:], generator=gen, dtype=torch.float64) - 0.5
    keep = torch.zeros(shp[3], dtype=torch.float64)
    keep[g:g + levels] = 1.0
    dT = dth * keep * (p / P0) ** (R / cp)
    w = torch.zeros(shp, dtype=torch.float64)
    w[4] = p
    w[0] = p / (R * (T + dT))
    return w


def ensemble_members(capi, n=96, seed=2024):
    """Batched warm-bubble ensemble: grid sizes drawn by seed from
    {128,192,256,384,512} x {64,96,128}, 50 m cells, dry ideal gas with
    theta0 = 300 K, a Gaussian +0.5 K bubble of radius 250-2000 m at a seeded
    position, WENO5 + HLLC, implicit 3x3 column solve, reflecting x1,
    periodic x2.  Returns (configs, primitive states as numpy arrays)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    R, cp, g, p0, th0 = 287.0, 1004.0, 9.81, 1e5, 300.0
    cfgs, ws = [], []
    for _ in range(n):
        nx2 = int(rng.choice([128, 192, 256, 384, 512]))
        nx1 = int(rng.choice([64, 96, 128]))
        c = capi.make_config(nx1=nx1, nx2=nx2, nghost=3, x1min=0.0, x1max=50.0 * nx1,
                             x2min=0.0, x2max=50.0 * nx2, gammad=cp / (cp - R),
                             weight=8.31446261815324 / R, recon_vertical=("weno5", False, False),
                             riemann="hllc", grav1=-g, implicit_scheme=1,
                             bc=("reflecting", "reflecting", "periodic", "periodic",
                                 "periodic", "periodic"))
        shp = c.shape
        z = (np.arange(shp[3]) - 3 + 0.5) * 50.0
        x = (np.arange(shp[2]) - 3 + 0.5) * 50.0
        X, Z = np.meshgrid(x, z, indexing="ij")
        rad = rng.uniform(250.0, 2000.0)
        xc, zc = rng.uniform(0.0, 50.0 * nx2), rng.uniform(0.2, 0.5) * 50.0 * nx1
        T = th0 - g * Z / cp
        p = p0 * (T / th0) ** (cp / R)
        th = 0.5 * np.exp(-((X - xc) ** 2 + (Z - zc) ** 2) / rad ** 2)
        w = np.zeros(shp)
        w[4, 0] = p
        w[0, 0] = p / (R * T * (1.0 + th / th0))
        cfgs.append(c)
        ws.append(w)
    return cfgs, ws


def c5_sensitivity(capi, dev, nsteps=200, every=20, ndir=2, wind=(0.0, 0.0), richardson=False,
                   roll=0.0, eps=1e-6, f32=False, return_grad=False):
    """Adjoint sensitivity on the 3-D box 128 x 128 x 96 of 200 m with the
    capped background of the 3-D workload (FP64):
    loss = 0.5 sum (theta_final - theta_target)^2 after nsteps SSP-RK3 steps,
    gradient with respect to the initial theta perturbation of the lowest 20
    levels, through the hand-written adjoint with a checkpoint every `every`
    steps; checked against central differences (step 1e-6 K) along `ndir`
    seeded directions (Richardson-extrapolated from steps 1e-6 and 2e-6 K with
    `richardson`).  `wind` adds a uniform horizontal wind (x2, x3) in m/s: at
    rest every x2 / x3 face starts exactly on the |u| switch of the Roe solver,
    where central differences straddle a kink.  `roll` adds a convective roll
    in the x1-x2 plane of that vertical amplitude (m/s), v1 = roll sin(pi z /
    H) cos(2 pi x / L) with the matching divergence-free v2, so that the x1
    faces leave the upwind switch of LMARS as well.

    With `f32` the forward is the fused FP32 path (snapx_fused_rk3_steps) and
    the gradient the FP32-storage adjoint (FP32 tapes and adjoint fields,
    FP64 column solves; fused steps between the checkpoints); the target and
    the finite differences stay FP64."""
    import torch
    from snapy_linfel_amd import adjoint as A
    R, cp, p0 = 287.0, 1004.5, 1e5
    cfg = box3d_cfg(capi, n=(96, 128, 128))
    cfg.x1max, cfg.x2max, cfg.x3max = 200.0 * 96, 200.0 * 128, 200.0 * 128
    shp = cfg.shape
    z = (torch.arange(shp[3], dtype=torch.float64, device=dev) - 3 + 0.5) * 200.0
    T0, pb = capped_background(z)
    thb = T0 * (p0 / pb) ** (R / cp)
    gam = cfg.gammad

    H, Lx = 200.0 * 96, 200.0 * 128
    xs = (torch.arange(shp[2], dtype=torch.float64, device=dev) - 3 + 0.5) * 200.0
    X2, Z2 = torch.meshgrid(xs, z, indexing="ij")  # [n2, n1]
    v1r = roll * torch.sin(torch.pi * Z2 / H) * torch.cos(2 * torch.pi * X2 / Lx)
    v2r = -roll * (Lx / (2 * H)) * torch.cos(torch.pi * Z2 / H) * torch.sin(2 * torch.pi * X2 / Lx)

    def state(dth):
        th = thb.expand(shp[1:]).clone()
        th = th + torch.nn.functional.pad(dth, (3, shp[3] - 23, 3, 3, 3, 3))
        T = th * (pb / p0) ** (R / cp)
        rho = pb / (R * T)
        v1 = v1r.expand(shp[1:])
        v2 = wind[0] + v2r.expand(shp[1:])
        m1, m2, m3 = rho * v1, rho * v2, rho * wind[1]
        e = pb / (gam - 1.0) + 0.5 * rho * (v1 ** 2 + v2 ** 2 + wind[1] ** 2)
        return torch.stack([rho, m1, m2, m3, e])

    def theta(u):
        rho = u[0]
        ke = 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / rho
        p = (gam - 1.0) * (u[4] - ke)
        return p / (rho * R) * (p0 / p) ** (R / cp)

    gen = torch.Generator(device="cpu").manual_seed(1234)
    cshape = (shp[1] - 6, shp[2] - 6, 20)  # interior columns, lowest 20 levels
    dth0 = ((torch.rand(cshape, generator=gen, dtype=torch.float64) - 0.5)).to(dev)
    dtt = ((torch.rand(cshape, generator=gen, dtype=torch.float64) - 0.5)).to(dev)
    dt = 0.4
    ws_f = capi.rk3_workspace(cfg, capi.F64)
    w_f = torch.empty(shp, dtype=torch.float64, device=dev)

    def run(dth):
        u = state(dth)
        for _ in range(nsteps):
            capi.rk3_step(cfg, dt, u, w_f, ws_f)
        return u

    th_target = theta(run(dtt))

    def loss_of(u):
        return 0.5 * ((theta(u.double()) - th_target)[3:-3, 3:-3, 3:-3] ** 2).sum()

    if f32:
        ws32 = capi.fused_workspace(cfg)

        def run_fwd(dth):
            u = state(dth).float()
            capi.fused_rk3_steps(cfg, dt, nsteps, u, ws32)
            return u
    else:
        run_fwd = run
    run_fwd(dth0)  # warm-up (first launches, workspace)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L0 = float(loss_of(run_fwd(dth0)))
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    x = dth0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u_start = state(x).float() if f32 else state(x)
    y = A.integrate(u_start, cfg, dt, nsteps, checkpoint_every=every)
    loss = loss_of(y)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    (grad,) = torch.autograd.grad(loss, x)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    t_back = time.perf_counter() - t1
    rels, fds, errs = [], [], []
    for k in range(ndir):
        d = torch.randn(cshape, generator=gen, dtype=torch.float64).to(dev)

        def central(eps):
            return (float(loss_of(run(dth0 + eps * d))) - float(loss_of(run(dth0 - eps * d)))) \
                / (2 * eps)
        # the flow crosses branch points of the schemes (upwind switches,
        # |u|) at some faces within any finite step, so differences converge
        # at first order over many steps: a small step in FP64
        fd = (4 * central(eps) - central(2 * eps)) / 3 if richardson else central(eps)
        ad = float((grad * d).sum())
        rels.append(abs(ad - fd) / abs(fd))
        fds.append(fd)
        errs.append(abs(ad - fd))
        if ndir > 4 and (k + 1) % 8 == 0:
            print(f"c5 finite differences: {k + 1}/{ndir} directions, max rel err "
                  f"{max(rels):.3e}", file=sys.stderr, flush=True)
    out = {"steps": nsteps, "checkpoint_every": every, "wind_m_s": list(wind), "roll_m_s": roll,
           "precision": ("FP32 fused forward; FP32-storage adjoint with FP64 column solves"
                         if f32 else "FP64 plain forward and adjoint"),
           "fd": (f"richardson {eps:g}/{2 * eps:g}" if richardson else f"central {eps:g}"),
           "loss": L0,
           "forward_s": round(t_fwd, 3), "forward_plus_adjoint_s": round(t_grad, 3),
           "gradient_over_forward": round(t_grad / t_fwd, 3),
           "backward_s": round(t_back, 3), "backward_over_forward": round(t_back / t_fwd, 3),
           "replay": "each segment recorded once (tapes of stage inputs, primitives, "
                     "tendencies, column blocks); the last segment recorded in the forward",
           "fd_directions": ndir,
           "fd_rel_err_max": max(rels) if rels else None,
           "fd_rel_err_median": sorted(rels)[len(rels) // 2] if rels else None}
    if rels:
        # g.d over random Gaussian directions is itself Gaussian, so among
        # many directions a few have a small directional derivative by
        # cancellation, and their relative error is large at the same
        # absolute error: also the error against the RMS directional
        # derivative of the sampled directions
        rms = (sum(f * f for f in fds) / len(fds)) ** 0.5
        out["fd_err_over_rms_fd_max"] = max(errs) / rms
        out["fd_err_over_rms_fd_median"] = sorted(errs)[len(errs) // 2] / rms
        out["fd_min_abs_fd_over_rms"] = min(abs(f) for f in fds) / rms
        out["fd_abs_fd_over_rms_at_max_rel_err"] = abs(fds[rels.index(max(rels))]) / rms
    return (out, grad) if return_grad else out


# VALU issue ceiling of MI355X: 256 CUs x 4 SIMDs, a wave64 VALU instruction
# every 4 cycles per SIMD (16 lanes per cycle; FP64 and non-packed FP32 FMA
# alike: 78.6 TFLOP/s = 256 x 4 x 16 x 2 flop x 2.4 GHz)
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4
# VALU floating-point peak without packed FP32 (disabled in these kernels):
# FP64 and FP32 FMA alike, 78.6 TFLOP/s
VALU_FLOP_PEAK = 78.6e12


def measured_flops(kernel, workload):
    """(FP32, FP64) VALU operations per launch from the SQ_INSTS_VALU_* counters
    in profiles/fused_kernels_pmc.md, or None."""
    path = os.path.join(ROOT, "profiles", "fused_kernels_pmc.md")
    key = f"flops {kernel} {workload}:"
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    w = line[len(key):].split()
                    return float(w[0]), float(w[2])
    except OSError:
        pass
    return None


def measured_traffic(kernel, workload, what="traffic"):
    """HBM bytes per launch of a kernel (what = "traffic"), as recorded from
    the FETCH_SIZE and WRITE_SIZE counters in profiles/fused_kernels_pmc.md,
    or its VALU wave-instructions per launch (what = "valu"); None if absent."""
    path = os.path.join(ROOT, "profiles", "fused_kernels_pmc.md")
    key = f"{what} {kernel} {workload}:"
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    return float(line[len(key):].split()[0])
    except OSError:
        pass
    return None


def errors(got, ref, interior, vel_scale=300.0):
    """Per conserved variable, against the field's own norm: Linf of the
    difference over max |ref| and the relative L2 norm; for the momenta also
    the Linf over max(max |ref|, max(rho) * vel_scale)."""
    import numpy as np
    out = {}
    names = ["rho", "mom1", "mom2", "mom3", "E"]
    rho_max = np.abs(ref[0][interior]).max()
    for v, name in enumerate(names):
        g, r = got[v][interior], ref[v][interior]
        if np.abs(r).max() == 0:
            continue
        d = np.abs(g - r).max()
        e = {"linf_rel": float(d / np.abs(r).max()),
             "l2_rel": float(np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-300))}
        if v in (1, 2, 3):
            e["linf_rel_vs_rho_max_x_%g_m_s" % vel_scale] = float(
                d / max(np.abs(r).max(), rho_max * vel_scale))
        out[name] = e
    return out


def max_rel(got, ref, interior):
    """Per conserved variable, max |got - ref| / max |ref| over the interior."""
    import numpy as np
    names = ["rho", "mom1", "mom2", "mom3", "E"]
    return {n: float(np.abs(got[v][interior] - ref[v][interior]).max()
                     / max(np.abs(ref[v][interior]).max(), 1e-300))
            for v, n in enumerate(names)}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores():
    try:
        return min(len(os.sched_getaffinity(0)), 16)
    except AttributeError:
        return min(os.cpu_count() or 1, 16)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                        "")
    except OSError:
        return ""


def cpu_baseline(cfg, u0, dt, steps=2, sub=64):
    """The FP64 C oracle (OpenMP over cells and columns) on a `steps`-step
    window of the workload, on the host cores this process may use (at most
    16); for scale, one step on one thread of a sub x sub column patch of the
    same block (periodic horizontally, so the same per-cell work).  Returns
    (record, state after the window)."""
    import oracle as O
    oc = O.OcConfig.from_buffer_copy(cfg)
    cores = host_cores()
    O.set_threads(cores)
    u = u0.copy()
    t0 = time.perf_counter()
    for _ in range(steps):
        u, _ = O.rk3_step(oc, dt, u)
    tn = time.perf_counter() - t0
    ps = O.OcConfig.from_buffer_copy(cfg)
    ps.nx2, ps.nx3 = sub, sub
    ps.x2max = cfg.x2min + (cfg.x2max - cfg.x2min) * sub / cfg.nx2
    ps.x3max = cfg.x3min + (cfg.x3max - cfg.x3min) * sub / cfg.nx3
    patch = O.apply_bc(ps, u0[:, :sub + 6, :sub + 6, :].copy(), O.CONSERVED)
    O.set_threads(1)
    t0 = time.perf_counter()
    O.rk3_step(ps, dt, patch)
    t1 = time.perf_counter() - t0
    n = cfg.n_interior
    rec = {"value": round(n * steps / tn, 1), "unit": "cell-updates/s", "cores": cores,
           "cpu_model": cpu_model(), "cpus_online": os.cpu_count(), "kind": "port",
           "value_1thread": round(sub * sub * cfg.nx1 / t1, 1),
           "sample": f"FP64 C oracle (restatement of the reference CPU path): the headline "
                     f"workload's first {steps} RK3 steps on {cores} threads ({tn:.1f} s); "
                     f"value_1thread: one step of a {sub}x{sub}x{cfg.nx1} column patch of it on "
                     f"1 thread ({t1:.1f} s)"}
    return rec, u


def kernel_roofline(name, avg_ms, bytes_, workload):
    """Roofline record of one kernel: algorithmic bytes per launch over its
    average device time, with the counter traffic and VALU rates recorded in
    profiles/fused_kernels_pmc.md for this kernel and workload."""
    achieved = bytes_ / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic(name, workload),
            "kernel": name, "avg_launch_ms": round(avg_ms, 5), "bytes_per_launch": bytes_}
    valu = measured_traffic(name, workload, "valu")
    if valu:
        roof["valu_wave_instructions_per_launch"] = valu
        roof["valu_issue_frac"] = round(valu / (avg_ms * 1e-3) / VALU_ISSUE_PEAK, 4)
    fl = measured_flops(name, workload)
    if fl:
        roof["valu_flops_per_launch"] = {"f32": fl[0], "f64": fl[1]}
        roof["valu_tflops"] = round((fl[0] + fl[1]) / (avg_ms * 1e-3) / 1e12, 3)
        roof["valu_flop_frac"] = round((fl[0] + fl[1]) / (avg_ms * 1e-3) / VALU_FLOP_PEAK, 4)
    return roof


def copy_rate_gbs(dev):
    """Attainable HBM rate: a 2 GiB device-to-device copy (read + write)."""
    import torch
    src = torch.empty(1 << 29, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    return 2 * src.numel() * 4 * 10 / (e0.elapsed_time(e1) * 1e-3) / 1e9


def fused_vs_plain(capi, cfg, u0, dt, K, W, dev, timing_steps=3, bgcols=None):
    """Times K fused FP32 steps and K plain FP64 steps from the same state
    (after W warm-up steps on copies); per-kernel device times of the fused
    step from HIP events over `timing_steps` more steps.  With bgcols =
    (rho_bar, p_bar) per x1 level, also the FP32 perturbation path about that
    FP64 background (split from the FP64 state, joined back to FP64)."""
    import torch
    out = {}
    if bgcols is not None:
        bg = capi.Background(cfg, *bgcols)
        wsb = capi.fused_workspace(cfg)
        up = bg.split(torch.from_numpy(u0).to(dev, torch.float64))
        bg.steps(dt, W, up.clone(), wsb)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bg.steps(dt, K, up, wsb)
        torch.cuda.synchronize()
        out["tbg"] = time.perf_counter() - t0
        tim = capi.Timing()
        bg.steps(dt, timing_steps, up.clone(), wsb, timing=tim)
        torch.cuda.synchronize()
        out["bg_rhs_ms"] = tim.rhs_ms / max(tim.rhs_launches, 1)
        out["bg_col_ms"] = tim.column_ms / max(tim.column_launches, 1)
        ubg = torch.empty(up.shape, dtype=torch.float64, device=dev)
        bg.join(up, ubg)
        out["ubg"] = ubg
        del wsb, up, bg
    u32 = torch.from_numpy(u0).to(dev, torch.float32)
    ws32 = capi.fused_workspace(cfg)
    capi.fused_rk3_steps(cfg, dt, W, u32.clone(), ws32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    capi.fused_rk3_steps(cfg, dt, K, u32, ws32)
    torch.cuda.synchronize()
    t32 = time.perf_counter() - t0
    tim = capi.Timing()
    capi.fused_rk3_steps(cfg, dt, timing_steps, u32.clone(), ws32, timing=tim)
    torch.cuda.synchronize()
    del ws32
    u64 = torch.from_numpy(u0).to(dev, torch.float64)
    w64 = torch.empty_like(u64)
    ws64 = capi.rk3_workspace(cfg, capi.F64)
    uw = u64.clone()
    for _ in range(max(W, 1)):
        capi.rk3_step(cfg, dt, uw, w64, ws64)
    del uw
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        capi.rk3_step(cfg, dt, u64, w64, ws64)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    del ws64, w64
    out.update({"u32": u32, "u64": u64, "t32": t32, "t64": t64,
                "rhs_ms": tim.rhs_ms / max(tim.rhs_launches, 1),
                "col_ms": tim.column_ms / max(tim.column_launches, 1),
                "launches": int(tim.rhs_launches + tim.column_launches + tim.bc_launches)})
    return out


def adjoint_ratio(capi, cfg, u0, dt, dev, n=3):
    """FP64 adjoint step (it recomputes its three stages) over the FP64 plain
    forward step, on the same state."""
    import torch
    u = torch.from_numpy(u0).to(dev, torch.float64)
    w = torch.empty_like(u)
    ws = capi.rk3_workspace(cfg, capi.F64)
    capi.rk3_step(cfg, dt, u.clone(), w, ws)
    torch.cuda.synchronize()
    uf = u.clone()
    t0 = time.perf_counter()
    for _ in range(n):
        capi.rk3_step(cfg, dt, uf, w, ws)
    torch.cuda.synchronize()
    tf = (time.perf_counter() - t0) / n
    del ws, uf, w
    wsa = capi.adjoint_workspace(cfg)
    lam = torch.randn(u.shape, dtype=torch.float64,
                      generator=torch.Generator().manual_seed(5)).to(dev)
    capi.adjoint_rk3_step(cfg, dt, u, lam.clone(), wsa)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        capi.adjoint_rk3_step(cfg, dt, u, lam, wsa)
    torch.cuda.synchronize()
    ta = (time.perf_counter() - t0) / n
    finite = bool(torch.isfinite(lam).all())
    del wsa, lam, u
    return {"forward_ms_per_step": round(tf * 1e3, 3), "adjoint_ms_per_step": round(ta * 1e3, 3),
            "adjoint_over_forward": round(ta / tf, 3), "finite": finite}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps of the 3-D workload")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle check")
    ap.add_argument("--no-straka", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-sph", action="store_true", help="skip the spherical-polar sector")
    ap.add_argument("--no-adjoint", action="store_true", help="skip every adjoint timing")
    ap.add_argument("--no-c5", action="store_true", help="skip the 200-step sensitivity run")
    ap.add_argument("--c5-fd-directions", type=int, default=64,
                    help="central-difference directions of the kink-free C5 check")
    ap.add_argument("--c5-directions", type=int, default=2,
                    help="random directions of the finite-difference check of the gradient")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the package has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from snapy_linfel_amd import capi

    dev = torch.device("cuda:0")
    K, W = args.steps, args.warmup
    extra = {}

    # ---------------- headline: 3-D box 256 x 256 x 128 ----------------
    cfg = box3d_cfg(capi)
    dt = 0.2  # horizontal acoustic CFL 0.7 on 100 m cells (0.5 s blows up, oracle too)
    u0 = conserved_on_device(capi, cfg, box3d_prim(cfg).numpy(), dev).cpu().numpy()
    z = (torch.arange(cfg.shape[3], dtype=torch.float64) - 3 + 0.5) * 100.0
    Tb, pb = capped_background(z, gamma=cfg.gammad)
    bgcols = ((pb / (287.0 * Tb)).numpy(), pb.numpy())
    I = (slice(3, -3), slice(3, -3), slice(3, -3))  # interior of one field [n3, n2, n1]
    log(f"3-D box 256x256x128: {K} fused FP32 steps (perturbation path and total state) and "
        f"{K} plain FP64 steps")
    r = fused_vs_plain(capi, cfg, u0, dt, K, W, dev, bgcols=bgcols)
    # headline: the FP32 fused step about the FP64 hydrostatic background
    cus32 = cfg.n_interior * K / r["tbg"]
    cus32_total = cfg.n_interior * K / r["t32"]
    cus64 = cfg.n_interior * K / r["t64"]
    got, ref = r["ubg"].cpu().numpy(), r["u64"].cpu().numpy()
    got_tot = r["u32"].double().cpu().numpy()
    err = errors(got, ref, I)
    err_tot = errors(got_tot, ref, I)
    m0 = u0[0][I].sum()
    mass_drift = float(abs(got[0][I].sum() - m0) / m0)
    mass_drift_tot = float(abs(got_tot[0][I].sum() - m0) / m0)
    mass_drift64 = float(abs(ref[0][I].sum() - m0) / m0)
    finite = bool(np.isfinite(got).all())
    del r["u32"], r["u64"], r["ubg"], got, ref, got_tot

    # roofline of the dominant kernel: K1 (fused RHS, 3-D marching tiles)
    # reads the stage state and writes the tendency, 2 NV 4 B per interior
    # cell; K2 (column solve + stage combine, two rows per lane) reads the
    # tendency, the stage state and u0 and writes the state, 4 NV 4 B
    n = cfg.n_interior
    k1 = ("fused_rhs_march_kernel<bg>", r["bg_rhs_ms"], 2 * NV * 4 * n)
    k2 = ("fused_column_pcr2_kernel<bg>", r["bg_col_ms"], 4 * NV * 4 * n)
    dom = k1 if k1[1] >= k2[1] else k2
    roof = kernel_roofline(*dom, "box256x256x128")
    other = k2 if dom is k1 else k1
    roof["other_kernel"] = {"kernel": other[0], "avg_launch_ms": round(other[1], 5),
                            "hbm_frac": round(other[2] / (other[1] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                              5)}
    ovr = kernel_roofline(*other, "box256x256x128")
    for key in ("traffic", "valu_issue_frac", "valu_flop_frac"):
        if key in ovr:
            roof["other_kernel"][key] = ovr[key]
    copy_gbs = copy_rate_gbs(dev)
    roof["copy_measured_GBs"] = round(copy_gbs, 1)
    roof["frac_of_copy"] = round(roof["achieved"] / copy_gbs, 5)

    # ---------------- oracle check and CPU baseline ----------------
    cpu, oracle_check = None, None
    if not args.no_cpu:
        log("CPU baseline: FP64 oracle, 2 steps of the 3-D box")
        cpu, u_or = cpu_baseline(cfg, u0, dt, steps=2)
        a = torch.from_numpy(u0).to(dev, torch.float32)
        capi.fused_rk3_steps(cfg, dt, 2, a)
        bgo = capi.Background(cfg, *bgcols)
        up = bgo.split(torch.from_numpy(u0).to(dev, torch.float64))
        bgo.steps(dt, 2, up)
        ab = torch.empty(up.shape, dtype=torch.float64, device=dev)
        bgo.join(up, ab)
        b = torch.from_numpy(u0).to(dev, torch.float64)
        wb = torch.empty_like(b)
        wsb = capi.rk3_workspace(cfg, capi.F64)
        for _ in range(2):
            capi.rk3_step(cfg, dt, b, wb, wsb)
        torch.cuda.synchronize()
        oracle_check = {"steps": 2,
                        "fp64_plain_max_rel": max_rel(b.cpu().numpy(), u_or, I),
                        "fp32_fused_background_max_rel": max_rel(ab.cpu().numpy(), u_or, I),
                        "fp32_fused_total_state_max_rel": max_rel(a.double().cpu().numpy(), u_or,
                                                                  I)}
        del a, b, wb, wsb, u_or, ab, up, bgo

    # ---------------- adjoint / forward on the headline block ----------------
    adj = None
    if not args.no_adjoint:
        log("adjoint / forward, FP64, 3-D box 256x256x128")
        adj = adjoint_ratio(capi, cfg, u0, dt, dev)
    del u0

    # ---------------- Straka 512 x 64 ----------------
    if not args.no_straka:
        log("Straka 512x64: 1000 fused FP32 and 1000 plain FP64 steps")
        cs = straka_cfg(capi)
        us0 = conserved_on_device(capi, cs, straka_prim(cs), dev).cpu().numpy()
        Is = (slice(None), slice(3, -3), slice(3, -3))
        wbg = straka_prim(cs, dT=0.0)
        rs = fused_vs_plain(capi, cs, us0, 0.1, 1000, 10, dev, timing_steps=100,
                            bgcols=(wbg[0, 0, 0].copy(), wbg[4, 0, 0].copy()))
        g32, g64 = rs["u32"].double().cpu().numpy(), rs["u64"].cpu().numpy()
        gbg = rs["ubg"].cpu().numpy()
        ms0 = us0[0][Is].sum()
        n2 = cs.n_interior
        ks1 = ("fused_rhs_kernel", rs["rhs_ms"], 2 * NV * 4 * n2)
        ks2 = ("fused_column_pcr2_kernel", rs["col_ms"], 4 * NV * 4 * n2)
        doms = ks1 if ks1[1] >= ks2[1] else ks2
        extra["straka_2d_512x64_fp32"] = {
            "cell_updates_per_s": round(n2 * 1000 / rs["t32"], 1), "steps": 1000, "dt_s": 0.1,
            "ms_per_step": round(rs["t32"] / 1000 * 1e3, 5),
            "fp64_plain_cell_updates_per_s": round(n2 * 1000 / rs["t64"], 1),
            "fused_over_fp64_plain": round(rs["t64"] / rs["t32"], 3),
            "fused_rhs_avg_ms": round(rs["rhs_ms"], 5),
            "fused_column_avg_ms": round(rs["col_ms"], 5),
            "roofline": kernel_roofline(*doms, "straka512x64"),
            "error_fp32_vs_fp64_after_steps": errors(g32, g64, Is),
            "mass_drift_fp32": float(abs(g32[0][Is].sum() - ms0) / ms0),
            "mass_drift_fp64": float(abs(g64[0][Is].sum() - ms0) / ms0),
            "reference_convention_value": round(
                cs.shape[1] * cs.shape[2] * cs.shape[3] * 1000 * 3 / rs["t32"], 1),
            "background_path": {
                "cell_updates_per_s": round(n2 * 1000 / rs["tbg"], 1),
                "ms_per_step": round(rs["tbg"] / 1000 * 1e3, 5),
                "fused_over_fp64_plain": round(rs["t64"] / rs["tbg"], 3),
                "fused_rhs_avg_ms": round(rs["bg_rhs_ms"], 5),
                "fused_column_avg_ms": round(rs["bg_col_ms"], 5),
                "error_fp32_vs_fp64_after_steps": errors(gbg, g64, Is),
                "mass_drift_fp32": float(abs(gbg[0][Is].sum() - ms0) / ms0)},
        }
        if not args.no_adjoint:
            extra["straka_2d_512x64_fp32"]["adjoint_fp64"] = adjoint_ratio(capi, cs, us0, 0.1,
                                                                          dev, n=20)
        del rs, g32, g64, gbg

    # ---- spherical-polar shell sector (64 radial x 512 polar cells):
    #      FP32 fused against FP64 plain on the same 100 steps ----
    if not args.no_sph:
        log("spherical-polar sector 64x512: 100 fused FP32 and 100 plain FP64 steps")
        cs = sector_cfg(capi)
        us0 = conserved_on_device(capi, cs, sector_prim(cs), dev).cpu().numpy()
        rs = fused_vs_plain(capi, cs, us0, 0.002, 100, 2, dev)
        Is = (slice(None), slice(3, -3), slice(3, -3))
        extra["spherical_sector_64x512_fp32"] = {
            "cell_updates_per_s": round(cs.n_interior * 100 / rs["t32"], 1), "steps": 100,
            "dt": 0.002, "ms_per_step": round(rs["t32"] / 100 * 1e3, 4),
            "fp64_plain_cell_updates_per_s": round(cs.n_interior * 100 / rs["t64"], 1),
            "fused_over_fp64_plain": round(rs["t64"] / rs["t32"], 3),
            "fused_rhs_avg_ms": round(rs["rhs_ms"], 5),
            "fused_column_avg_ms": round(rs["col_ms"], 5),
            "error_fp32_vs_fp64_after_steps": errors(rs["u32"].double().cpu().numpy(),
                                                     rs["u64"].cpu().numpy(), Is, vel_scale=1.0),
            "finite": bool(torch.isfinite(rs["u32"]).all()),
        }
        del rs

    # ---- batched ensemble (96 members of uneven size), autotuned tiles,
    #      one HIP-graph launch of the configuration's 500 steps ----
    if not args.no_ensemble:
        log("ensemble: 96 members, autotune, 500 steps in one graph launch")
        cfgs_e, ws_e = ensemble_members(capi)
        parts = []
        for c, w in zip(cfgs_e, ws_e):
            parts.append(conserved_on_device(capi, c, w, dev).float().reshape(-1))
        ue = torch.cat(parts)
        del parts
        dte = 0.1
        b = capi.Batch(cfgs_e, dte)
        tile, rhs_ms = b.autotune(ue, dte, trials=3)
        steps_e = 500
        u_keep = ue.clone()
        g = b.graph(ue, steps_e)
        g.launch()
        torch.cuda.synchronize()
        ue.copy_(u_keep)
        del u_keep
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.launch()
        torch.cuda.synchronize()
        te = time.perf_counter() - t0
        ncells_e = sum(c.n_interior for c in cfgs_e)
        extra["ensemble_96_fp32_graph"] = {
            "members": len(cfgs_e), "cells": int(ncells_e), "steps": steps_e,
            "cell_updates_per_s": round(ncells_e * steps_e / te, 1),
            "ms_per_step": round(te / steps_e * 1e3, 4),
            "autotuned_tile": list(tile), "rhs_ms_at_tile": round(rhs_ms, 4),
            "finite": bool(torch.isfinite(ue).all()),
        }
        del g, b, ue

    # ---- adjoint sensitivity: 3-D box 128 x 128 x 96 of 200 m, FP64 ----
    if not args.no_adjoint and not args.no_c5:
        log("adjoint sensitivity, 128x128x96, 200 steps (FP64)")
        c5, g64 = c5_sensitivity(capi, dev, ndir=args.c5_directions, return_grad=True)
        extra["c5_sensitivity_box3d_128x128x96_fp64"] = c5
        log("adjoint sensitivity, 128x128x96, 200 steps (FP32 fused forward, FP32 adjoint)")
        c5f, g32 = c5_sensitivity(capi, dev, ndir=0, f32=True, return_grad=True)
        dg = (g32 - g64).double()
        c5f["gradient_vs_fp64_adjoint"] = {
            "l2_rel": float(dg.norm() / g64.norm()),
            "linf_rel": float(dg.abs().max() / g64.abs().max())}
        c5f["backward_over_fp64_backward"] = round(c5f["backward_s"] / c5["backward_s"], 3)
        extra["c5_sensitivity_box3d_128x128x96_fp32"] = c5f
        del g64, g32, dg
        # the same gradient check off the kinks of the state at rest: a
        # uniform horizontal wind (Roe |u| at the x2 / x3 faces) and a
        # convective roll (LMARS upwind switch at the x1 faces)
        extra["c5_fd_check_wind_and_roll"] = c5_sensitivity(
            capi, dev, ndir=args.c5_fd_directions, wind=(2.0, 1.5), roll=1.0)

    line = {
        "metric": "cell-updates/s per RK3 step, FP32 fused vs FP64 plain at stated err; "
                  "adjoint/fwd",
        "value": round(cus32, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(r["tbg"] / K * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (perturbation about an FP64 hydrostatic background; FP64 implicit "
                 "column solve)",
        "data": "synthetic (capped dry-adiabat background with seeded theta noise)",
        "config": {"workload": "box3d_256x256x128_100m_weno5_lmars_x1_roe_x2x3_vic3_dt0.2",
                   "cells": [256, 256, 128], "dt_s": dt, "steps": K},
        # the reference's own counter (src/mesh/meshblock.cpp:938-956): cells
        # including ghosts x steps x stages / seconds
        "reference_convention_value": round(
            cfg.shape[1] * cfg.shape[2] * cfg.shape[3] * K * 3 / r["tbg"], 1),
        "fp64_plain": {"value": round(cus64, 1), "ms_per_step": round(r["t64"] / K * 1e3, 5)},
        "fp32_fused_over_fp64_plain": round(cus32 / cus64, 3),
        "error_fp32_vs_fp64_after_steps": err,
        "mass_drift_fp32": mass_drift,
        "mass_drift_fp64": mass_drift64,
        # the same fused step on the FP32 total state (no background)
        "fp32_total_state": {"value": round(cus32_total, 1),
                             "ms_per_step": round(r["t32"] / K * 1e3, 5),
                             "fused_rhs_avg_ms": round(r["rhs_ms"], 5),
                             "fused_column_avg_ms": round(r["col_ms"], 5),
                             "error_fp32_vs_fp64_after_steps": err_tot,
                             "mass_drift_fp32": mass_drift_tot},
        "finite": finite,
        "oracle_check": oracle_check,
        "adjoint_over_forward": adj["adjoint_over_forward"] if adj else None,
        "adjoint_note": "FP64 adjoint step (recomputes its 3 stages) / FP64 plain forward step "
                        "on the headline block",
        "adjoint": adj,
        "fused_rhs_avg_ms": round(r["bg_rhs_ms"], 5),
        "fused_column_avg_ms": round(r["bg_col_ms"], 5),
        "step_hbm_frac": round(cus32 * 160 / (HBM_PEAK_GBS * 1e9), 6),
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
