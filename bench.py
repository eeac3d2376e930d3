"""Benchmark: cell-updates/s per SSP-RK3 step on one MI355X.

Workload (the configuration the metric is quoted on): 2-D Straka density
current, 512 x 64 cells (x2 horizontal, periodic, 25.6 km; x1 vertical,
reflecting, 6.4 km), dry ideal gas cp = 1004, R = 287, theta0 = 300 K,
-15 K cosine bubble, gravity 9.81, WENO5 (shock: false) + LMARS, vertically
implicit 3x3 solve, SSP-RK3 with dt = 0.1 s.  Synthetic initial state (the
problem set-up of examples/straka.cpp), inputs resident in HBM.

Reported:
  value        FP32 fused path, interior cells x steps / wall seconds
  fp64_plain   the FP64 plain (module-by-module) GPU path on the same steps
  error        per-field L2 / Linf relative error of FP32 fused vs FP64 plain
               after the timed steps, and the FP32 mass drift
  roofline     dominant kernel of the fused step, algorithmic bytes per
               launch / its average device time (HIP events on the stream)
  cpu_baseline the FP64 CPU oracle on a bounded sample of the workload
  step_hbm_frac  160 B per cell-step (FP32) x CU/s / 8 TB/s
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
                   roll=0.0, eps=1e-6):
    """Adjoint sensitivity on the 3-D box 128 x 128 x 96 of 200 m (FP64):
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
    faces leave the upwind switch of LMARS as well."""
    import torch
    from snapy_linfel_amd import adjoint as A
    R, cp, g0, p0 = 287.0, 1004.5, 9.81, 1e5
    cfg = box3d_cfg(capi, n=(96, 128, 128))
    cfg.x1max, cfg.x2max, cfg.x3max = 200.0 * 96, 200.0 * 128, 200.0 * 128
    shp = cfg.shape
    z = (torch.arange(shp[3], dtype=torch.float64, device=dev) - 3 + 0.5) * 200.0
    T0 = 250.0 - g0 * z / cp
    pb = p0 * (T0 / 250.0) ** (cp / R)
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
        return 0.5 * ((theta(u) - th_target)[3:-3, 3:-3, 3:-3] ** 2).sum()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L0 = float(loss_of(run(dth0)))
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    x = dth0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y = A.integrate(state(x), cfg, dt, nsteps, checkpoint_every=every)
    loss = loss_of(y)
    (grad,) = torch.autograd.grad(loss, x)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    rels = []
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
        if ndir > 4 and (k + 1) % 8 == 0:
            print(f"c5 finite differences: {k + 1}/{ndir} directions, max rel err "
                  f"{max(rels):.3e}", file=sys.stderr, flush=True)
    return {"steps": nsteps, "checkpoint_every": every, "wind_m_s": list(wind), "roll_m_s": roll,
            "fd": (f"richardson {eps:g}/{2 * eps:g}" if richardson else f"central {eps:g}"),
            "loss": L0,
            "forward_s": round(t_fwd, 3), "forward_plus_adjoint_s": round(t_grad, 3),
            "gradient_over_forward": round(t_grad / t_fwd, 3),
            "fd_directions": ndir, "fd_rel_err_max": max(rels),
            "fd_rel_err_median": sorted(rels)[len(rels) // 2]}


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
    """Per conserved variable: Linf of the difference over max |ref| (momenta:
    at least max(rho) * vel_scale), and the relative L2 norm."""
    import numpy as np
    out = {}
    names = ["rho", "mom1", "mom2", "mom3", "E"]
    rho_max = np.abs(ref[0][interior]).max()
    for v, name in enumerate(names):
        g, r = got[v][interior], ref[v][interior]
        if np.abs(r).max() == 0:
            continue
        scale = max(np.abs(r).max(), rho_max * vel_scale if v in (1, 2, 3) else 0.0)
        out[name] = {
            "linf_rel": float(np.abs(g - r).max() / scale),
            "l2_rel": float(np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-300)),
        }
    return out


def cpu_baseline(steps_cap=400, seconds=10.0):
    """The FP64 C oracle (OpenMP over cells and columns) on the host cores this
    process may use (at most 16), and on one thread for scale."""
    import oracle as O
    oc = O.straka_config(nx2=512, nx1=64, implicit_scheme=1)
    w, u0 = O.initialize(oc, O.straka_init(oc))

    def run(threads, secs, cap):
        O.set_threads(threads)
        u = u0
        t0 = time.perf_counter()
        n = 0
        while n < cap and time.perf_counter() - t0 < secs:
            u, _ = O.rk3_step(oc, 0.1, u)
            n += 1
        return n, time.perf_counter() - t0

    try:
        cores = min(len(os.sched_getaffinity(0)), 16)
    except AttributeError:
        cores = min(os.cpu_count() or 1, 16)
    n1, t1 = run(1, seconds * 0.5, steps_cap)
    nn, tn = run(cores, seconds, steps_cap)
    O.set_threads(1)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                         "")
    except OSError:
        pass
    return {"value": round(512 * 64 * nn / tn, 1), "unit": "cell-updates/s", "cores": cores,
            "cpu_model": model, "cpus_online": os.cpu_count(),
            "kind": "port", "value_1thread": round(512 * 64 * n1 / t1, 1),
            "sample": f"FP64 C oracle, Straka 512x64, RK3 steps from the initial state: "
                      f"{nn} steps on {cores} threads ({tn:.1f} s), "
                      f"{n1} steps on 1 thread ({t1:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-sph", action="store_true", help="skip the spherical-polar sector")
    ap.add_argument("--no-c5", action="store_true", help="skip the 200-step sensitivity run")
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
    cfg = straka_cfg(capi)
    u0 = conserved_on_device(capi, cfg, straka_prim(cfg), dev).cpu().numpy()
    I = (slice(None), slice(3, -3), slice(3, -3))
    dt = 0.1
    K, W = args.steps, args.warmup

    # ---- FP32 fused: warm-up, then K timed steps (launches only) ----
    u32 = torch.from_numpy(u0).to(dev, torch.float32)
    ws32 = capi.fused_workspace(cfg)
    capi.fused_rk3_steps(cfg, dt, W, u32.clone(), ws32)
    torch.cuda.synchronize()
    u32.copy_(torch.from_numpy(u0).to(dev, torch.float32))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    capi.fused_rk3_steps(cfg, dt, K, u32, ws32)
    torch.cuda.synchronize()
    t32 = time.perf_counter() - t0
    cus32 = cfg.n_interior * K / t32

    # ---- per-kernel device times (HIP events around every launch) ----
    tim = capi.Timing()
    ut = torch.from_numpy(u0).to(dev, torch.float32)
    nk = min(K, 200)
    capi.fused_rk3_steps(cfg, dt, nk, ut, ws32, timing=tim)
    torch.cuda.synchronize()
    rhs_avg = tim.rhs_ms / max(tim.rhs_launches, 1)
    col_avg = tim.column_ms / max(tim.column_launches, 1)

    # ---- FP64 plain path on the same steps ----
    u64 = torch.from_numpy(u0).to(dev, torch.float64)
    w64 = torch.empty_like(u64)
    ws64 = capi.rk3_workspace(cfg, capi.F64)
    for _ in range(W):
        capi.rk3_step(cfg, dt, u64.clone(), w64, ws64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        capi.rk3_step(cfg, dt, u64, w64, ws64)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    cus64 = cfg.n_interior * K / t64

    # ---- adjoint of the FP64 step (stage recomputation included) ----
    ws_adj = capi.adjoint_workspace(cfg)
    lam = torch.randn(u64.shape, dtype=torch.float64,
                      generator=torch.Generator().manual_seed(5)).to(dev)
    na = max(1, min(K, 50))
    capi.adjoint_rk3_step(cfg, dt, u64, lam.clone(), ws_adj)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(na):
        capi.adjoint_rk3_step(cfg, dt, u64, lam, ws_adj)
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / na
    adj_ratio = t_adj / (t64 / K)
    del ws_adj, lam

    got = u32.double().cpu().numpy()
    ref = u64.cpu().numpy()
    err = errors(got, ref, I)
    m0 = u0[0][I].sum()
    mass_drift = float(abs(got[0][I].sum() - m0) / m0)
    mass_drift64 = float(abs(ref[0][I].sum() - m0) / m0)

    # ---- roofline of the dominant kernel ----
    # K1 (fused RHS) algorithmic bytes: read the stage state + write the
    # tendency, 2 * NV * 4 B per interior cell; K2 (column solve + combine):
    # read tendency, stage state, u0 and write the state, 4 * NV * 4 B.
    n = cfg.n_interior
    k1_bytes, k2_bytes = 2 * NV * 4 * n, 4 * NV * 4 * n
    if rhs_avg >= col_avg:
        name, avg_ms, bytes_ = "fused_rhs_kernel", rhs_avg, k1_bytes
    else:
        name, avg_ms, bytes_ = "fused_column_pcr_kernel<3>", col_avg, k2_bytes
    achieved = bytes_ / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic(name, "straka512x64"),
            "kernel": name, "avg_launch_ms": round(avg_ms, 5),
            "bytes_per_launch": bytes_}
    valu = measured_traffic(name, "straka512x64", "valu")
    if valu:
        roof["valu_wave_instructions_per_launch"] = valu
        roof["valu_issue_frac"] = round(valu / (avg_ms * 1e-3) / VALU_ISSUE_PEAK, 4)
    fl = measured_flops(name, "straka512x64")
    if fl:
        roof["valu_flops_per_launch"] = {"f32": fl[0], "f64": fl[1]}
        roof["valu_tflops"] = round((fl[0] + fl[1]) / (avg_ms * 1e-3) / 1e12, 3)
        roof["valu_flop_frac"] = round((fl[0] + fl[1]) / (avg_ms * 1e-3) / VALU_FLOP_PEAK, 4)

    # attainable HBM rate: a 2 GiB device-to-device copy (read + write bytes)
    src = torch.empty(1 << 29, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * src.numel() * 4 * 10 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    roof["copy_measured_GBs"] = round(copy_gbs, 1)
    roof["frac_of_copy"] = round(achieved / copy_gbs, 5)

    # the floor a launch of this size meets on this GPU: a 64-element kernel,
    # and a device copy moving the dominant kernel's algorithmic bytes (half
    # read, half written); 100 launches captured in a HIP graph and replayed,
    # GPU time per launch
    def per_launch_us(fn, n):
        fn()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                for _ in range(100):
                    fn()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(n // 100):
            g.replay()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / (n // 100 * 100) * 1e3
    tiny = torch.zeros(64, device=dev)
    csrc = torch.ones(bytes_ // 8, dtype=torch.float32, device=dev)
    cdst = torch.empty_like(csrc)
    t_tiny = per_launch_us(lambda: tiny.add_(1.0), 2000)
    t_copy = per_launch_us(lambda: cdst.copy_(csrc), 2000)
    roof["floor"] = {"tiny_kernel_us": round(t_tiny, 3), "copy_same_bytes_us": round(t_copy, 3),
                     "copy_same_bytes_GBs": round(bytes_ / (t_copy * 1e-6) / 1e9, 1),
                     "kernel_over_copy_same_bytes": round(avg_ms * 1e3 / t_copy, 2)}
    del tiny, csrc, cdst

    extra = {"fused_rhs_avg_ms": round(rhs_avg, 5), "fused_column_avg_ms": round(col_avg, 5),
             "timed_launches": int(tim.rhs_launches + tim.column_launches + tim.bc_launches)}

    # ---- 3-D box (256 x 256 x 128, configuration-3 shape) secondary line ----
    if not args.no_3d:
        c3 = box3d_cfg(capi)
        shp = c3.shape
        z = (torch.arange(shp[3], dtype=torch.float64) - 3 + 0.5) * 100.0
        T = 250.0 - 9.81 * z / 1004.5
        p = 1e5 * (T / 250.0) ** (1004.5 / 287.0)
        w3 = torch.zeros(shp, dtype=torch.float64)
        w3[4] = p
        g = torch.Generator().manual_seed(1234)
        th = (torch.rand(shp[1:], generator=g, dtype=torch.float64) - 0.5)
        th[..., 23:] = 0.0
        w3[0] = p / (287.0 * (T + th))
        wd = w3.to(dev)
        ud = torch.empty_like(wd)
        capi.prim2cons(c3, wd, ud)
        capi.apply_bc(c3, ud)
        u3 = ud.float()
        ws3 = capi.fused_workspace(c3)
        dt3 = 0.2  # horizontal acoustic CFL 0.7 on 100 m cells
        capi.fused_rk3_steps(c3, dt3, 2, u3.clone(), ws3)
        torch.cuda.synchronize()
        k3 = 100  # the configuration's 100 steps
        t0 = time.perf_counter()
        capi.fused_rk3_steps(c3, dt3, k3, u3, ws3)
        torch.cuda.synchronize()
        t3 = time.perf_counter() - t0
        tim3 = capi.Timing()
        capi.fused_rk3_steps(c3, dt3, 3, u3, ws3, timing=tim3)
        cus3 = c3.n_interior * k3 / t3
        r3 = tim3.rhs_ms / max(tim3.rhs_launches, 1)
        c3ms = tim3.column_ms / max(tim3.column_launches, 1)
        extra["box3d_256x256x128_fp32"] = {
            "cell_updates_per_s": round(cus3, 1), "steps": k3, "dt_s": dt3,
            "finite": bool(torch.isfinite(u3).all()),
            "ms_per_step": round(t3 / k3 * 1e3, 4),
            "step_hbm_frac": round(cus3 * 160 / (HBM_PEAK_GBS * 1e9), 4),
            "fused_rhs_avg_ms": round(r3, 4), "fused_column_avg_ms": round(c3ms, 4),
            "rhs_hbm_frac": round(2 * NV * 4 * c3.n_interior / (r3 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                  4),
        }
        for key, kname, ms in (("rhs", "fused_rhs_march_kernel", r3),
                               ("column", "fused_column_pcr2_kernel", c3ms)):
            valu = measured_traffic(kname, "box256x256x128", "valu")
            if valu:
                extra["box3d_256x256x128_fp32"][f"{key}_valu_issue_frac"] = round(
                    valu / (ms * 1e-3) / VALU_ISSUE_PEAK, 4)
            fl = measured_flops(kname, "box256x256x128")
            if fl:
                extra["box3d_256x256x128_fp32"][f"{key}_valu_tflops"] = round(
                    (fl[0] + fl[1]) / (ms * 1e-3) / 1e12, 3)
                extra["box3d_256x256x128_fp32"][f"{key}_valu_flop_frac"] = round(
                    (fl[0] + fl[1]) / (ms * 1e-3) / VALU_FLOP_PEAK, 4)

    # ---- spherical-polar shell sector (64 radial x 512 polar cells):
    #      FP32 fused against FP64 plain on the same 100 steps ----
    if not args.no_sph:
        cs = sector_cfg(capi)
        us64 = conserved_on_device(capi, cs, sector_prim(cs), dev)
        us32 = us64.float()
        s0 = us64.clone()
        dts, ks = 0.002, 100  # x2 acoustic CFL ~0.8 at r = 1
        wss = capi.fused_workspace(cs)
        capi.fused_rk3_steps(cs, dts, 2, us32.clone(), wss)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        capi.fused_rk3_steps(cs, dts, ks, us32, wss)
        torch.cuda.synchronize()
        ts32 = time.perf_counter() - t0
        tims = capi.Timing()
        capi.fused_rk3_steps(cs, dts, 3, us32.clone(), wss, timing=tims)
        ws64s = capi.rk3_workspace(cs, capi.F64)
        wt = torch.empty_like(us64)
        capi.rk3_step(cs, dts, us64.clone(), wt, ws64s)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(ks):
            capi.rk3_step(cs, dts, us64, wt, ws64s)
        torch.cuda.synchronize()
        ts64 = time.perf_counter() - t0
        Is = (slice(None), slice(3, -3), slice(3, -3))
        es = errors(us32.double().cpu().numpy(), us64.cpu().numpy(), Is, vel_scale=1.0)
        extra["spherical_sector_64x512_fp32"] = {
            "cell_updates_per_s": round(cs.n_interior * ks / ts32, 1), "steps": ks, "dt": dts,
            "ms_per_step": round(ts32 / ks * 1e3, 4),
            "fp64_plain_cell_updates_per_s": round(cs.n_interior * ks / ts64, 1),
            "fused_over_fp64_plain": round(ts64 / ts32, 3),
            "fused_rhs_avg_ms": round(tims.rhs_ms / max(tims.rhs_launches, 1), 5),
            "fused_column_avg_ms": round(tims.column_ms / max(tims.column_launches, 1), 5),
            "error_fp32_vs_fp64_after_steps": es,
            "finite": bool(torch.isfinite(us32).all()),
        }
        del wss, ws64s, us32, us64, s0, wt

    # ---- batched ensemble (96 members of uneven size), autotuned tiles,
    #      HIP-graph replays ----
    if not args.no_ensemble:
        cfgs_e, ws_e = ensemble_members(capi)
        parts = []
        for c, w in zip(cfgs_e, ws_e):
            wd = torch.from_numpy(w).to(dev)
            ud = torch.empty_like(wd)
            capi.prim2cons(c, wd, ud)
            capi.apply_bc(c, ud)
            parts.append(ud.float().reshape(-1))
        ue = torch.cat(parts)
        del parts
        dte = 0.1
        b = capi.Batch(cfgs_e, dte)
        tile, rhs_ms = b.autotune(ue, dte, trials=3)
        # the configuration's 500 steps as one graph launch (a warm-up launch
        # first, on a copy of the state)
        steps_per_graph, ngraph = 500, 1
        u_keep = ue.clone()
        g = b.graph(ue, steps_per_graph)
        g.launch()
        torch.cuda.synchronize()
        ue.copy_(u_keep)
        del u_keep
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(ngraph):
            g.launch()
        torch.cuda.synchronize()
        te = time.perf_counter() - t0
        ncells_e = sum(c.n_interior for c in cfgs_e)
        extra["ensemble_96_fp32_graph"] = {
            "members": len(cfgs_e), "cells": int(ncells_e), "steps": steps_per_graph * ngraph,
            "cell_updates_per_s": round(ncells_e * steps_per_graph * ngraph / te, 1),
            "ms_per_step": round(te / (steps_per_graph * ngraph) * 1e3, 4),
            "autotuned_tile": list(tile), "rhs_ms_at_tile": round(rhs_ms, 4),
            "finite": bool(torch.isfinite(ue).all()),
        }
        del g, b, ue

    # ---- adjoint at scale: 3-D box 128 x 128 x 96 of 200 m, FP64 ----
    if not args.no_3d:
        ca = box3d_cfg(capi, n=(96, 128, 128))
        ca.x1max, ca.x2max, ca.x3max = 200.0 * 96, 200.0 * 128, 200.0 * 128
        shp = ca.shape
        z = (torch.arange(shp[3], dtype=torch.float64) - 3 + 0.5) * 200.0
        T = 250.0 - 9.81 * z / 1004.5
        p = 1e5 * (T / 250.0) ** (1004.5 / 287.0)
        wa = torch.zeros(shp, dtype=torch.float64)
        wa[4] = p
        g2 = torch.Generator().manual_seed(99)
        th = (torch.rand(shp[1:], generator=g2, dtype=torch.float64) - 0.5)
        th[..., 23:] = 0.0
        wa[0] = p / (287.0 * (T + th))
        wad = wa.to(dev)
        ua = torch.empty_like(wad)
        capi.prim2cons(ca, wad, ua)
        capi.apply_bc(ca, ua)
        dta = 0.4
        wsf = capi.rk3_workspace(ca, capi.F64)
        wwa = torch.empty_like(ua)
        capi.rk3_step(ca, dta, ua.clone(), wwa, wsf)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        uf = ua.clone()
        for _ in range(3):
            capi.rk3_step(ca, dta, uf, wwa, wsf)
        torch.cuda.synchronize()
        tf = (time.perf_counter() - t0) / 3
        del wsf
        wsa = capi.adjoint_workspace(ca)
        lam_a = torch.randn(ua.shape, dtype=torch.float64, generator=g2).to(dev)
        capi.adjoint_rk3_step(ca, dta, ua, lam_a.clone(), wsa)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            capi.adjoint_rk3_step(ca, dta, ua, lam_a, wsa)
        torch.cuda.synchronize()
        tadj = (time.perf_counter() - t0) / 3
        if not args.no_c5:
            extra["c5_sensitivity_box3d_128x128x96_fp64"] = c5_sensitivity(
                capi, dev, ndir=args.c5_directions)
            # the same gradient check off the kinks of the state at rest: a
            # uniform horizontal wind (Roe |u| at the x2 / x3 faces) and a
            # convective roll (LMARS upwind switch at the x1 faces)
            extra["c5_fd_check_wind_and_roll"] = c5_sensitivity(
                capi, dev, ndir=8, wind=(2.0, 1.5), roll=1.0)
        extra["adjoint_box3d_128x128x96_fp64"] = {
            "forward_ms_per_step": round(tf * 1e3, 3), "adjoint_ms_per_step": round(tadj * 1e3, 3),
            "adjoint_over_forward": round(tadj / tf, 3),
            "finite": bool(torch.isfinite(lam_a).all()),
        }
        del wsa, lam_a

    cpu = None if args.no_cpu else cpu_baseline()

    line = {
        "metric": "cell-updates/s per RK3 step, FP32 fused vs FP64 plain at stated err; "
                  "adjoint/fwd",
        "value": round(cus32, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(t32 / K * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 implicit column solve)",
        "data": "synthetic (Straka initial state of examples/straka.cpp)",
        "config": {"workload": "straka_2d_512x64_weno5_lmars_vic3_dt0.1",
                   "cells": [512, 64], "dt_s": dt, "steps": K},
        # the reference's own counter (src/mesh/meshblock.cpp:938-956): cells
        # including ghosts x steps x stages / seconds
        "reference_convention_value": round(
            cfg.shape[1] * cfg.shape[2] * cfg.shape[3] * K * 3 / t32, 1),
        "fp64_plain": {"value": round(cus64, 1), "ms_per_step": round(t64 / K * 1e3, 5)},
        "fp32_fused_over_fp64_plain": round(cus32 / cus64, 3),
        "error_fp32_vs_fp64_after_steps": err,
        "mass_drift_fp32": mass_drift,
        "mass_drift_fp64": mass_drift64,
        "adjoint_over_forward": round(adj_ratio, 3),
        "adjoint_note": "FP64 adjoint step (recomputes its 3 stages) / FP64 plain forward step",
        "adjoint_ms_per_step": round(t_adj * 1e3, 4),
        "step_hbm_frac": round(cus32 * 160 / (HBM_PEAK_GBS * 1e9), 6),
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
