"""Differentiable PyTorch restatement of one SSP-RK3 step of the hot path.

TEST INFRASTRUCTURE ONLY.  Like the C oracle next to it, this module is a
checker: only tests import it.  Nothing in the product calls it.

It restates, in torch operations that autograd can differentiate (FP64,
CPU), the same step the C oracle (``snap_oracle.c``) and the FP64 plain GPU
path compute, so that the hand-written adjoint can be checked against
libtorch autograd.  The reference's own CPU kernels write through raw
pointers and cannot be differentiated; its MPS fallbacks are pure-torch
formulations of the same pointwise kernels, and this restatement follows the
same algebra:

* cons2prim            src/eos/ideal_gas_impl.h:15-35 (ideal_gas_cons2prim_mps,
                       src/eos/eos_dispatch.cpp:38-55)
* WENO5 / CP5          src/recon/interp_impl.h:22-31, 73-122; tables of
                       src/recon/weno5.cpp:10-24, src/recon/cp5.cpp:12-53
                       (call_weno5_mps / call_poly_mps, src/recon/recon_dispatch.cpp)
* reconstruct rules    src/recon/reconstruct.cpp:75-154 (shock / CP5 split)
* ghost refill, LR     src/hydro/hydro.cpp:217-269
* LMARS / HLLC / Roe   src/riemann/lmars_impl.h:16-78, hllc_impl.h:16-130,
                       roe_impl.h:23-211 (the *_mps formulations)
* divergence, gravity  src/coord/coordinate.cpp:352-396,
                       src/forcing/const_gravity.cpp:26-45
* VIC                  src/implicit/vic_assemble_partial_impl.h:28-111,
                       vic_assemble_full_impl.h:15-106, forward_sweep_impl.h:20-109,
                       vic_redistribute_impl.h:26-106 (block Thomas in torch
                       ops: there is no torch formulation in the reference)
* boundaries, RK3      src/bc/bc_func.cpp:7-63, src/mesh/meshblock.cpp:577-786

Pinned by ``tests/test_torch_oracle.py`` against the C oracle (same step to
1e-12).  Supported: dry ideal gas, WENO5 / CP5 (scale off), LMARS / HLLC /
Roe, gravity along x1, implicit schemes 0 / 1 / 9, reflecting / outflow /
periodic boundaries, nghost 3.
"""
from __future__ import annotations

import torch

IDN, IVX, IVY, IVZ, IPR = 0, 1, 2, 3, 4
INTERP_CP5, INTERP_WENO5 = 3, 5
RIEMANN_ROE, RIEMANN_LMARS, RIEMANN_HLLC = 0, 1, 2
BC_REFLECTING, BC_OUTFLOW, BC_PERIODIC = 0, 1, 2


def _shape(c):
    ng = c.nghost
    return (c.nx3 + 2 * ng if c.nx3 > 1 else 1, c.nx2 + 2 * ng if c.nx2 > 1 else 1,
            c.nx1 + 2 * ng if c.nx1 > 1 else 1)


def _dx(c):
    return ((c.x1max - c.x1min) / c.nx1, (c.x2max - c.x2min) / c.nx2,
            (c.x3max - c.x3min) / c.nx3)


def cons2prim(c, u):
    rho = u[IDN]
    v = u[IVX:IPR] / rho
    ke = 0.5 * ((v[0] * u[IVX] + v[1] * u[IVY]) + v[2] * u[IVZ])
    p = (c.gammad - 1.0) * (u[IPR] - ke)
    return torch.cat([rho[None], v, p[None]], 0)


def prim2cons(c, w):
    rho = w[IDN]
    m = w[IVX:IPR] * rho
    e = w[IPR] / (c.gammad - 1.0) + 0.5 * rho * ((w[IVX] ** 2 + w[IVY] ** 2) + w[IVZ] ** 2)
    return torch.cat([rho[None], m, e[None]], 0)


def interior(c):
    """Index of the interior cells of [5, n3, n2, n1] (ghosts along axes > 1)."""
    ng = c.nghost
    return (slice(None),) + tuple(slice(ng, -ng) if n > 1 else slice(None)
                                  for n in _shape(c))


def check_redo(c, u):
    """Verdict bits (1 density, 2 pressure, 4 non-finite) of
    MeshBlockImpl::check_redo (src/mesh/meshblock.cpp:974-1028), limiter off:
    NaN-propagating tensor minima over the interior compared with zero; the
    pressure test runs only when the conserved-energy screen trips."""
    I = interior(c)
    ui = u[I]
    rho = bool(ui[IDN].min() <= 0.0)
    pres = False
    if bool(ui[IPR].min() <= 0.0):
        pres = bool(cons2prim(c, u)[I][IPR].min() <= 0.0)
    nonfinite = not bool(torch.isfinite(ui).all())
    return int(rho) | (int(pres) << 1) | (int(nonfinite) << 2)


# ------------------------------------------------------------- boundaries

def _axis_bc(x, ax, ng, bin_, bout, normal_var, flip):
    """One axis of the boundary pass on [nvar, n3, n2, n1] (tensor axis ax)."""
    n = x.shape[ax]
    inner = x.narrow(ax, ng, n - 2 * ng)
    sgn = torch.ones(x.shape[0], dtype=x.dtype)
    if flip:
        sgn[normal_var] = -1.0
    sgn = sgn.view(-1, 1, 1, 1)

    def ghosts(kind, lower):
        if kind == BC_REFLECTING:
            src = x.narrow(ax, ng, ng) if lower else x.narrow(ax, n - 2 * ng, ng)
            return src.flip(ax) * sgn
        if kind == BC_OUTFLOW:
            edge = x.narrow(ax, ng, 1) if lower else x.narrow(ax, n - ng - 1, 1)
            return torch.cat([edge] * ng, ax)
        if kind == BC_PERIODIC:
            return x.narrow(ax, n - 2 * ng, ng) if lower else x.narrow(ax, ng, ng)
        return x.narrow(ax, 0, ng) if lower else x.narrow(ax, n - ng, ng)

    return torch.cat([ghosts(bin_, True), inner, ghosts(bout, False)], ax)


def apply_bc(c, x, flip=True):
    """x1, then x2, then x3 (bc_kernel order); flip the normal momentum."""
    ng = c.nghost
    for a, (ax, nx) in enumerate(((3, c.nx1), (2, c.nx2), (1, c.nx3))):
        if nx <= 1 or (c.bc[2 * a] < 0 and c.bc[2 * a + 1] < 0):
            continue
        x = _axis_bc(x, ax, ng, c.bc[2 * a], c.bc[2 * a + 1], IVX + a, flip)
    return x


def _refill_ghosts(c, w):
    """Isentropic hydrostatic refill of the x1 ghosts of rho and p."""
    ng = c.nghost
    n1 = w.shape[3]
    gamma = c.gammad
    ae = (gamma - 1.0) / gamma
    grav = -c.grav1
    dz = _dx(c)[0]
    rho, pre = w[IDN], w[IPR]
    out_r, out_p = [None] * n1, [None] * n1
    for side in (0, 1):
        is_ = ng if side == 0 else n1 - ng - 1
        K = pre[..., is_] / rho[..., is_] ** gamma
        prev = pre[..., is_]
        for n in range(ng):
            h = prev ** ae + ae * grav * dz / K ** (1.0 / gamma) if side == 0 else \
                prev ** ae - ae * grav * dz / K ** (1.0 / gamma)
            pg = h ** (1.0 / ae)
            g = is_ - n - 1 if side == 0 else is_ + n + 1
            out_p[g] = pg
            out_r[g] = (pg / K) ** (1.0 / gamma)
            prev = pg
    for i in range(ng, n1 - ng):
        out_r[i], out_p[i] = rho[..., i], pre[..., i]
    r = torch.stack(out_r, -1)
    p = torch.stack(out_p, -1)
    return torch.cat([r[None], w[IVX:IPR], p[None]], 0)


# ----------------------------------------------------------- interpolants

def _weno5_lo(f0, f1, f2, f3, f4):
    p0 = (-1. / 6. * f0 + 5. / 6. * f1) + 1. / 3. * f2
    p1 = (1. / 3. * f1 + 5. / 6. * f2) + -1. / 6. * f3
    p2 = (11. / 6. * f2 + -7. / 6. * f3) + 1. / 3. * f4
    b0 = 13. / 12. * ((f0 - 2. * f1) + f2) ** 2 + .25 * ((f0 - 4. * f1) + 3. * f2) ** 2
    b1 = 13. / 12. * ((f1 - 2. * f2) + f3) ** 2 + .25 * (-f1 + f3) ** 2
    b2 = 13. / 12. * ((f2 - 2. * f3) + f4) ** 2 + .25 * ((3. * f2 - 4. * f3) + f4) ** 2
    a0, a1, a2 = .3 / (b0 + 1e-6) ** 2, .6 / (b1 + 1e-6) ** 2, .1 / (b2 + 1e-6) ** 2
    return ((a0 * p0 + a1 * p1) + a2 * p2) / ((a0 + a1) + a2)


def _weno5_hi(f0, f1, f2, f3, f4):
    p0 = (1. / 3. * f2 + 5. / 6. * f3) + -1. / 6. * f4
    p1 = (-1. / 6. * f1 + 5. / 6. * f2) + 1. / 3. * f3
    p2 = (1. / 3. * f0 + -7. / 6. * f1) + 11. / 6. * f2
    b0 = 13. / 12. * ((f2 - 2. * f3) + f4) ** 2 + .25 * ((3. * f2 - 4. * f3) + f4) ** 2
    b1 = 13. / 12. * ((f1 - 2. * f2) + f3) ** 2 + .25 * (f1 - f3) ** 2
    b2 = 13. / 12. * ((f0 - 2. * f1) + f2) ** 2 + .25 * ((f0 - 4. * f1) + 3. * f2) ** 2
    a0, a1, a2 = .3 / (b0 + 1e-6) ** 2, .6 / (b1 + 1e-6) ** 2, .1 / (b2 + 1e-6) ** 2
    return ((a0 * p0 + a1 * p1) + a2 * p2) / ((a0 + a1) + a2)


def _cp5_lo(f0, f1, f2, f3, f4):
    return (((-1. / 20. * f0 + 9. / 20. * f1) + 47. / 60. * f2) + -13. / 60. * f3) + 1. / 30. * f4


def _cp5_hi(f0, f1, f2, f3, f4):
    return (((1. / 30. * f0 + -13. / 60. * f1) + 47. / 60. * f2) + 9. / 20. * f3) + -1. / 20. * f4


def _face_states(c, w, ax, r):
    """Left / right states at faces f = ng .. n-ng along tensor axis ax (the
    faces whose fluxes reach interior cells)."""
    ng = c.nghost
    n = w.shape[ax]
    nf = n - 2 * ng + 1
    W = [w.narrow(ax, ng - 3 + k, nf) for k in range(6)]  # cells f-3+k
    typ, shock = c.recon_type[r], c.recon_shock[r]
    if typ not in (INTERP_WENO5, INTERP_CP5) or c.recon_scale[r]:
        raise NotImplementedError("torch oracle: WENO5 / CP5 without scale only")
    wl, wr = [], []
    for v in range(5):
        t = typ if (shock or v == IDN) else INTERP_CP5
        lo, hi = (_weno5_lo, _weno5_hi) if t == INTERP_WENO5 else (_cp5_lo, _cp5_hi)
        wr.append(lo(*[W[k][v] for k in range(1, 6)]))  # lower face of cell f
        wl.append(hi(*[W[k][v] for k in range(0, 5)]))  # upper face of cell f-1
    return torch.stack(wl, 0), torch.stack(wr, 0)


# ----------------------------------------------------------------- Riemann

def _frame(w, dim):
    """(rho, vn, vt1, vt2, p) with the reference's ivx = IPR - dim slots."""
    ivx = IPR - dim
    ivy = IVX + ((ivx - IVX) + 1) % 3
    ivz = IVX + ((ivx - IVX) + 2) % 3
    return ivx, ivy, ivz


def _lmars(c, wl, wr, dim):
    g = c.gammad
    ivx, ivy, ivz = _frame(wl, dim)
    L = [wl[IDN], wl[ivx], wl[ivy], wl[ivz], wl[IPR]]
    R = [wr[IDN], wr[ivx], wr[ivy], wr[ivz], wr[IPR]]
    hl = (wl[IPR] / (g - 1.)) / wl[IDN] + 0.5 * ((L[1] ** 2 + L[2] ** 2) + L[3] ** 2) + L[4] / L[0]
    hr = (wr[IPR] / (g - 1.)) / wr[IDN] + 0.5 * ((R[1] ** 2 + R[2] ** 2) + R[3] ** 2) + R[4] / R[0]
    rhobar = 0.5 * (L[0] + R[0])
    cbar = torch.sqrt(0.5 * g * (L[4] + R[4]) / rhobar)
    pbar = 0.5 * (L[4] + R[4]) + 0.5 * (rhobar * cbar) * (L[1] - R[1])
    ubar = 0.5 * (L[1] + R[1]) + 0.5 / (rhobar * cbar) * (L[4] - R[4])
    up = ubar > 0.0
    S = [torch.where(up, a, b) for a, b in zip(L, R)]
    h = torch.where(up, hl, hr)
    f = [None] * 5
    f[IDN] = ubar * S[0]
    f[ivx] = ubar * S[0] * S[1] + pbar
    f[ivy] = ubar * S[0] * S[2]
    f[ivz] = ubar * S[0] * S[3]
    f[IPR] = ubar * S[0] * h
    return torch.stack(f, 0)


def _hllc(c, wl, wr, dim):
    g = c.gammad
    tiny = 1.0e-10
    ivx, ivy, ivz = _frame(wl, dim)
    el = wl[IPR] / (g - 1.) + 0.5 * wl[IDN] * ((wl[IVX] ** 2 + wl[IVY] ** 2) + wl[IVZ] ** 2)
    er = wr[IPR] / (g - 1.) + 0.5 * wr[IDN] * ((wr[IVX] ** 2 + wr[IVY] ** 2) + wr[IVZ] ** 2)
    cl = torch.sqrt(g * wl[IPR] / wl[IDN])
    cr = torch.sqrt(g * wr[IPR] / wr[IDN])
    rhoa = .5 * (wl[IDN] + wr[IDN])
    ca = .5 * (cl + cr)
    pmid = .5 * ((wl[IPR] + wr[IPR]) + (wl[ivx] - wr[ivx]) * rhoa * ca)
    # the sqrt argument is clamped on the untaken branch only (no NaN gradient)
    argl = torch.where(pmid <= wl[IPR], torch.ones_like(pmid),
                       1.0 + (g + 1) / (2 * g) * (pmid / wl[IPR] - 1.0))
    argr = torch.where(pmid <= wr[IPR], torch.ones_like(pmid),
                       1.0 + (g + 1) / (2 * g) * (pmid / wr[IPR] - 1.0))
    ql = torch.where(pmid <= wl[IPR], torch.ones_like(pmid), torch.sqrt(argl))
    qr = torch.where(pmid <= wr[IPR], torch.ones_like(pmid), torch.sqrt(argr))
    al = wl[ivx] - cl * ql
    ar = wr[ivx] + cr * qr
    bp = torch.where(ar > 0.0, ar, torch.full_like(ar, tiny))
    bm = torch.where(al < 0.0, al, torch.full_like(al, -tiny))
    vxl = wl[ivx] - al
    vxr = wr[ivx] - ar
    tl = wl[IPR] + vxl * wl[IDN] * wl[ivx]
    tr = wr[IPR] + vxr * wr[IDN] * wr[ivx]
    ml = wl[IDN] * vxl
    mr = -(wr[IDN] * vxr)
    am = (tl - tr) / (ml + mr)
    cp = (ml * tr + mr * tl) / (ml + mr)
    cp = torch.where(cp > 0.0, cp, torch.zeros_like(cp))
    vxl = wl[ivx] - bm
    vxr = wr[ivx] - bp
    fl, fr = [None] * 5, [None] * 5
    fl[IDN], fr[IDN] = wl[IDN] * vxl, wr[IDN] * vxr
    fl[ivx] = wl[IDN] * wl[ivx] * vxl + wl[IPR]
    fr[ivx] = wr[IDN] * wr[ivx] * vxr + wr[IPR]
    fl[ivy], fr[ivy] = wl[IDN] * wl[ivy] * vxl, wr[IDN] * wr[ivy] * vxr
    fl[ivz], fr[ivz] = wl[IDN] * wl[ivz] * vxl, wr[IDN] * wr[ivz] * vxr
    fl[IPR] = el * vxl + wl[IPR] * wl[ivx]
    fr[IPR] = er * vxr + wr[IPR] * wr[ivx]
    pos = am >= 0.0
    sl = torch.where(pos, am / (am - bm), torch.zeros_like(am))
    sr = torch.where(pos, torch.zeros_like(am), -am / (bp - am))
    sm = torch.where(pos, -bm / (am - bm), bp / (bp - am))
    f = [sl * fl[v] + sr * fr[v] for v in range(5)]
    f[ivx] = f[ivx] + sm * cp
    f[IPR] = f[IPR] + sm * cp * am
    return torch.stack(f, 0)


def _roe(c, wl, wr, dim):
    g = c.gammad
    tiny = 1.0e-10
    ivx, ivy, ivz = _frame(wl, dim)
    el = wl[IPR] / (g - 1.)
    er = wr[IPR] / (g - 1.)
    cl = torch.sqrt(g * wl[IPR] / wl[IDN])
    cr = torch.sqrt(g * wr[IPR] / wr[IDN])
    etl = el + 0.5 * wl[IDN] * ((wl[IVX] ** 2 + wl[IVY] ** 2) + wl[IVZ] ** 2)
    etr = er + 0.5 * wr[IDN] * ((wr[IVX] ** 2 + wr[IVY] ** 2) + wr[IVZ] ** 2)
    sdl, sdr = torch.sqrt(wl[IDN]), torch.sqrt(wr[IDN])
    isd = 1.0 / (sdl + sdr)
    rhobar = sdl * sdr
    v1 = (sdl * wl[ivx] + sdr * wr[ivx]) * isd
    v2 = (sdl * wl[ivy] + sdr * wr[ivy]) * isd
    v3 = (sdl * wl[ivz] + sdr * wr[ivz]) * isd
    hl = (etl + wl[IPR]) / wl[IDN]
    hr = (etr + wr[IPR]) / wr[IDN]
    h = (hl * sdl + hr * sdr) * isd
    rl0, rr0 = wl[IDN], wr[IDN]
    F = [rl0 * wl[ivx], wl[IDN] * wl[ivx] * wl[ivx] + wl[IPR], wl[IDN] * wl[ivx] * wl[ivy],
         wl[IDN] * wl[ivx] * wl[ivz], (etl + wl[IPR]) * wl[ivx]]
    G = [rr0 * wr[ivx], wr[IDN] * wr[ivx] * wr[ivx] + wr[IPR], wr[IDN] * wr[ivx] * wr[ivy],
         wr[IDN] * wr[ivx] * wr[ivz], (etr + wr[IPR]) * wr[ivx]]
    qbar0 = (rl0 / sdl + rr0 / sdr) * isd
    du1 = wr[IDN] * wr[ivx] - wl[IDN] * wl[ivx]
    du2 = wr[IDN] * wr[ivy] - wl[IDN] * wl[ivy]
    du3 = wr[IDN] * wr[ivz] - wl[IDN] * wl[ivz]
    du4 = etr - etl
    vsq = (v1 * v1 + v2 * v2) + v3 * v3
    gm1 = g - 1.0
    q = h - 0.5 * vsq
    qm = torch.where(q > tiny, q, torch.full_like(q, tiny))
    cs_sq = torch.where(gm1 * qm > tiny, gm1 * qm, torch.full_like(qm, tiny))
    cs = torch.sqrt(cs_sq)
    lam_m, lam_p = v1 - cs, v1 + cs
    spd_m, spd_0, spd_p = torch.abs(lam_m), torch.abs(v1), torch.abs(lam_p)
    du = wr[ivx] - wl[ivx]
    dv = wr[ivy] - wl[ivy]
    dw = wr[ivz] - wl[ivz]
    dp = wr[IPR] - wl[IPR]
    alpha_m = -0.5 * rhobar / cs * du + 0.5 * dp / cs_sq
    alpha_p = +0.5 * rhobar / cs * du + 0.5 * dp / cs_sq
    alpha_v, alpha_w = rhobar * dv, rhobar * dw
    alpha0 = (rr0 - rl0) - dp / cs_sq * qbar0
    rho_first = rl0 + alpha_m * qbar0
    llf = (rho_first < 0.0) | (rho_first + alpha0 < 0.0)
    f = [0.5 * (F[v] + G[v]) for v in range(5)]
    f0 = f[0] - 0.5 * ((spd_m * alpha_m * qbar0 + spd_0 * alpha0) + spd_p * alpha_p * qbar0)
    f1 = f[1] - 0.5 * ((spd_m * alpha_m * (v1 - cs) + spd_0 * (v1 * alpha0)) +
                       spd_p * alpha_p * (v1 + cs))
    f2 = f[2] - 0.5 * ((spd_m * alpha_m * v2 + spd_0 * (v2 * alpha0 + alpha_v)) +
                       spd_p * alpha_p * v2)
    f3 = f[3] - 0.5 * ((spd_m * alpha_m * v3 + spd_0 * (v3 * alpha0 + alpha_w)) +
                       spd_p * alpha_p * v3)
    f4 = f[4] - 0.5 * ((spd_m * alpha_m * (h - v1 * cs) +
                        spd_0 * ((rhobar * ((hr - hl) - v1 * du) + alpha0 * h) - dp)) +
                       spd_p * alpha_p * (h + v1 * cs))
    out = [f0, f1, f2, f3, f4]
    out = [torch.where(lam_m > 0.0, F[v], out[v]) for v in range(5)]
    out = [torch.where(lam_p < 0.0, G[v], out[v]) for v in range(5)]
    x, y = torch.abs(wl[ivx]) + cl, torch.abs(wr[ivx]) + cr
    a = 0.5 * torch.where(x > y, x, y)
    dd = [rr0 - rl0, du1, du2, du3, du4]
    out = [torch.where(llf, 0.5 * (F[v] + G[v]) - a * dd[v], out[v]) for v in range(5)]
    # slots -> variables
    res = [None] * 5
    res[IDN], res[ivx], res[ivy], res[ivz], res[IPR] = out
    return torch.stack(res, 0)


_SOLVERS = {RIEMANN_LMARS: _lmars, RIEMANN_HLLC: _hllc, RIEMANN_ROE: _roe}


# --------------------------------------------------------------------- VIC

def _flux_jac5(gm1, w):
    """5x5 flux Jacobian (flux_decomposition_impl.h:107-128) per cell."""
    rho, v1, v2, v3, p = w[IDN], w[IVX], w[IVY], w[IVZ], w[IPR]
    z, o = torch.zeros_like(rho), torch.ones_like(rho)
    s2 = (v1 * v1 + v2 * v2) + v3 * v3
    c1 = ((gm1 - 1) * s2 / 2 - (gm1 + 1) / gm1 * p / rho) * v1
    c2 = ((gm1 + 1) / gm1 * p / rho + s2 / 2) - gm1 * v1 * v1
    rows = [[z, o, z, z, z],
            [gm1 * s2 / 2 - v1 * v1, (2. - gm1) * v1, -gm1 * v2, -gm1 * v3, gm1 * o],
            [-v1 * v2, v2, v1, z, z],
            [-v1 * v3, v3, z, v1, z],
            [c1, c2, -gm1 * v2 * v1, -gm1 * v3 * v1, (gm1 + 1) * v1]]
    return torch.stack([torch.stack(r, -1) for r in rows], -2)


def _abs_jac5(gm1, wl, wr):
    """|A| = R |Lambda| R^-1 at the Roe average (flux_decomposition_impl.h:39-105)."""
    sdl, sdr = torch.sqrt(wl[IDN]), torch.sqrt(wr[IDN])
    isd = 1.0 / (sdl + sdr)
    r = sdl * sdr
    u = (sdl * wl[IVX] + sdr * wr[IVX]) * isd
    v = (sdl * wl[IVY] + sdr * wr[IVY]) * isd
    w = (sdl * wl[IVZ] + sdr * wr[IVZ]) * isd
    el = wl[IPR] / gm1 + 0.5 * wl[IDN] * ((wl[IVX] ** 2 + wl[IVY] ** 2) + wl[IVZ] ** 2)
    er = wr[IPR] / gm1 + 0.5 * wr[IDN] * ((wr[IVX] ** 2 + wr[IVY] ** 2) + wr[IVZ] ** 2)
    hbar = ((el + wl[IPR]) / sdl + (er + wr[IPR]) / sdr) * isd
    p = (hbar - 0.5 * ((u ** 2 + v ** 2) + w ** 2)) * gm1 / (gm1 + 1.) * r
    cs = torch.sqrt(p * (gm1 + 1.) / r)
    ke = 0.5 * ((u * u + v * v) + w * w)
    hp = (gm1 + 1.) / gm1 * p / r
    h = hp + ke
    z, o = torch.zeros_like(u), torch.ones_like(u)
    R = torch.stack([torch.stack(x, -1) for x in (
        [o, o, o, z, z], [u - cs, u, u + cs, z, z], [v, v, v, o, z], [w, w, w, z, o],
        [h - u * cs, ke, h + u * cs, v, w])], -2)
    tch, th = 2. * cs * hp, 2. * hp
    Ri = torch.stack([torch.stack(x, -1) for x in (
        [(cs * ke + u * hp) / tch, (-hp - cs * u) / tch, -v / th, -w / th, 1. / th],
        [(hp - ke) / hp, u / hp, v / hp, w / hp, -1. / hp],
        [(cs * ke - u * hp) / tch, (hp - cs * u) / tch, -v / th, -w / th, 1. / th],
        [-v, z, o, z, z], [-w, z, z, o, z])], -2)
    lam = torch.stack([torch.abs(u - cs), torch.abs(u), torch.abs(u + cs), torch.abs(u),
                       torch.abs(u)], -1)
    return (R * lam.unsqueeze(-2)) @ Ri


def _implicit(c, dt, du, w):
    """VIC correction of du on the interior columns (du, w: [5, n3, n2, n1])."""
    N = 5 if (c.implicit_scheme >> 3) & 1 else 3
    sub = [IDN, IVX, IVY, IVZ, IPR] if N == 5 else [IDN, IVX, IPR]
    ng, nl = c.nghost, c.nx1
    n3, n2, n1 = w.shape[1:]
    s3 = slice(ng, n3 - ng) if n3 > 1 else slice(0, 1)
    s2 = slice(ng, n2 - ng) if n2 > 1 else slice(0, 1)
    wc = w[:, s3, s2, ng - 1:ng + nl + 1].reshape(5, -1, nl + 2)  # [5, col, layer]
    dx1, dx2, dx3 = _dx(c)
    area, vol = dx3 * dx2, dx3 * (dx2 * dx1)
    gm1 = c.gammad - 1.0
    J = _flux_jac5(gm1, wc)                                        # [col, nl+2, 5, 5]
    A = _abs_jac5(gm1, wc[:, :, :-1], wc[:, :, 1:])                # faces [col, nl+1, 5, 5]
    idx = torch.tensor(sub)
    J = J[..., idx, :][..., :, idx]
    A = A[..., idx, :][..., :, idx]
    hv = 0.5 / vol
    Am, Ap = A[:, :-1], A[:, 1:]
    Jp, Jn = J[:, :-2], J[:, 2:]
    phi = torch.zeros(N, N, dtype=w.dtype)
    phi[1, 0] = c.grav1
    phi[N - 1, 1] = c.grav1
    eye = torch.eye(N, dtype=w.dtype)
    a = (Am * area + Ap * area) * hv - phi + eye / dt
    b = -(Am + Jp) * area * hv
    cc = -(Ap - Jn) * area * hv
    bnd = torch.ones(N, dtype=w.dtype)
    bnd[1] = -1.0
    a0 = a[:, 0] + b[:, 0] * bnd
    aL = a[:, -1] + cc[:, -1] * bnd
    if nl == 1:
        a = (a0 + cc[:, 0] * bnd).unsqueeze(1)
    else:
        a = torch.cat([a0[:, None], a[:, 1:-1], aL[:, None]], 1)
    duc = du[:, s3, s2, ng:ng + nl].reshape(5, -1, nl)
    rhs = (duc[idx] / dt).permute(1, 2, 0)                         # [col, nl, N]
    P, d = [None] * nl, [None] * nl
    for i in range(nl):
        ai, ri = a[:, i], rhs[:, i]
        if i > 0:
            ai = ai - b[:, i] @ P[i - 1]
            ri = ri - (b[:, i] @ d[i - 1].unsqueeze(-1)).squeeze(-1)
        inv = torch.linalg.inv(ai)
        P[i] = inv @ cc[:, i]
        d[i] = (inv @ ri.unsqueeze(-1)).squeeze(-1)
    delta = [None] * nl
    delta[nl - 1] = d[nl - 1]
    for i in range(nl - 2, -1, -1):
        delta[i] = d[i] - (P[i] @ delta[i + 1].unsqueeze(-1)).squeeze(-1)
    delta = torch.stack(delta, 1)                                   # [col, nl, N]
    rho = wc[IDN, :, 1:-1]
    ai_ = rho * vol
    S = (ai_ * ai_).sum(-1, keepdim=True)
    Bd = ((duc[IDN] - delta[..., 0]) * 1.0 * vol).sum(-1, keepdim=True)
    out = list(duc.unbind(0))
    out[IDN] = (duc[IDN] + (delta[..., 0] - duc[IDN])) + rho * ai_ * Bd / S
    for k, v in enumerate(sub[1:], start=1):
        out[v] = delta[..., k]
    new = torch.stack(out, 0).reshape(5, du[:, s3, s2].shape[1], du[:, s3, s2].shape[2], nl)
    # write back into the interior of du (differentiable)
    du = du.clone()
    du[:, s3, s2, ng:ng + nl] = new
    return du


# ------------------------------------------------------------- the step

def hydro_forward(c, dt, u):
    """du of one stage (HydroImpl::forward, src/hydro/hydro_forward.cpp:14-217)."""
    ng = c.nghost
    w = cons2prim(c, u)
    if c.nx1 > 1 and c.grav1 != 0.0:
        w = _refill_ghosts(c, w)
    n3, n2, n1 = w.shape[1:]
    s1 = slice(ng, n1 - ng)
    s2 = slice(ng, n2 - ng) if n2 > 1 else slice(0, 1)
    s3 = slice(ng, n3 - ng) if n3 > 1 else slice(0, 1)
    dx = _dx(c)
    div = 0.0
    for a, (ax, nx) in enumerate(((3, c.nx1), (2, c.nx2), (1, c.nx3))):
        if nx <= 1:
            continue
        r = 0 if a == 0 else 1
        wl, wr = _face_states(c, w, ax, r)
        if a == 0 and c.grav1 != 0.0:
            nf = wl.shape[ax]
            first = torch.zeros(nf, dtype=torch.bool)
            first[0] = True
            last = torch.zeros(nf, dtype=torch.bool)
            last[-1] = True
            for v in (IDN, IPR):
                wl = torch.cat([wl[:v], torch.where(first, wr[v], wl[v])[None], wl[v + 1:]], 0)
                wr = torch.cat([wr[:v], torch.where(last, wl[v], wr[v])[None], wr[v + 1:]], 0)
        F = _SOLVERS[c.riemann[a]](c, wl, wr, 3 - a)
        dF = (F.narrow(ax, 1, nx) - F.narrow(ax, 0, nx)) / dx[a]
        # keep the interior of the other axes
        sel = [slice(None), s3, s2, s1]
        sel[ax] = slice(None)
        div = div + dF[tuple(sel)]
    du = torch.zeros_like(u)
    du = du.clone()
    du[:, s3, s2, s1] = -dt * div
    rho = w[IDN]
    if c.grav1 != 0.0:
        add = torch.zeros_like(u)
        add[IVX] = dt * rho * c.grav1 * c.non_hydrostatic
        add[IPR] = dt * rho * w[IVX] * c.grav1 * c.non_hydrostatic
        du = du + add
    if c.implicit_scheme:
        du = _implicit(c, dt, du, w)
    return du


def rk3_step(c, dt, u):
    """One SSP-RK3 step: u_{s+1} = BC(w0 u0 + w1 u_s + w2 du(u_s))."""
    weights = ((0., 1., 1.), (3. / 4., 1. / 4., 1. / 4.), (1. / 3., 2. / 3., 2. / 3.))
    u0, us = u, u
    for w0, w1, w2 in weights:
        us = apply_bc(c, (w0 * u0 + w1 * us) + w2 * hydro_forward(c, dt, us))
    return us
