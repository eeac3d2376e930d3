"""Pins the oracle's LMARS / HLLC / Roe solvers.

The face-pressure closed forms and seeds replicate the reference's own tests
(tests/test_riemann.cpp:412-618); consistency, mirror symmetry and the Sod
problem against the exact Riemann solution (Toro, ch. 4) are fixed by the
mathematics.
"""
import numpy as np
import pytest
import torch

import oracle as O

GAMMA = 1.4
DIM1 = 3
SHAPE = (5, 12, 14, 16)  # nx1=10, nx2=8, nx3=6, nghost=3 (test_riemann.cpp:94-97)


def _states(seed):
    """tests/test_riemann.cpp:437-445 inputs, generated with the same torch
    CPU generator and seed."""
    torch.manual_seed(seed)
    wl = torch.rand(SHAPE, dtype=torch.float64)
    wr = torch.rand(SHAPE, dtype=torch.float64)
    wl[0].add_(1.0)
    wr[0].add_(1.0)
    wl[4].add_(1.0)
    wr[4].add_(1.0)
    wl.narrow(0, 1, 3).sub_(0.5).mul_(0.1)
    wr.narrow(0, 1, 3).sub_(0.5).mul_(0.1)
    return wl.numpy().copy(), wr.numpy().copy()


def test_hllc_face_pressure_closed_form():
    # tests/test_riemann.cpp:412-472, seed 77
    wl, wr = _states(77)
    cl = np.sqrt(GAMMA * wl[4] / wl[0])
    cr = np.sqrt(GAMMA * wr[4] / wr[0])
    rhoa = 0.5 * (wl[0] + wr[0])
    ca = 0.5 * (cl + cr)
    pmid = 0.5 * (wl[4] + wr[4] + (wl[1] - wr[1]) * rhoa * ca)
    ql = np.where(pmid <= wl[4], 1.0,
                  np.sqrt(1.0 + (GAMMA + 1) / (2 * GAMMA) * (pmid / wl[4] - 1.0)))
    qr = np.where(pmid <= wr[4], 1.0,
                  np.sqrt(1.0 + (GAMMA + 1) / (2 * GAMMA) * (pmid / wr[4] - 1.0)))
    al = wl[1] - cl * ql
    ar = wr[1] + cr * qr
    vxl = wl[1] - al
    vxr = wr[1] - ar
    tl = wl[4] + vxl * wl[0] * wl[1]
    tr = wr[4] + vxr * wr[0] * wr[1]
    ml = wl[0] * vxl
    mr = -(wr[0] * vxr)
    expected = np.maximum((ml * tr + mr * tl) / (ml + mr), 0.0)
    _, fp = O.riemann("hllc", wl, wr, DIM1)
    assert np.allclose(fp, expected, rtol=1e-10, atol=1e-10)


def test_lmars_face_pressure_closed_form():
    # tests/test_riemann.cpp:474-515, seed 78
    wl, wr = _states(78)
    rhobar = 0.5 * (wl[0] + wr[0])
    cbar = np.sqrt(GAMMA * 0.5 * (wl[4] + wr[4]) / rhobar)
    expected = 0.5 * (wl[4] + wr[4]) + 0.5 * (rhobar * cbar) * (wl[1] - wr[1])
    _, fp = O.riemann("lmars", wl, wr, DIM1)
    assert np.allclose(fp, expected, rtol=1e-10, atol=1e-10)


def test_roe_face_pressure_closed_form():
    # tests/test_riemann.cpp:517-570, seed 79
    wl, wr = _states(79)
    c = O.make_config(nx1=10, nx2=8, nx3=6)
    ul = O.prim2cons(c, wl)
    ur = O.prim2cons(c, wr)
    sdl, sdr = np.sqrt(wl[0]), np.sqrt(wr[0])
    isd = 1.0 / (sdl + sdr)
    hbar = ((ul[4] + wl[4]) / sdl + (ur[4] + wr[4]) / sdr) * isd
    vbar = (sdl * wl[1:4] + sdr * wr[1:4]) * isd
    cs = np.sqrt((GAMMA - 1.0) * (hbar - 0.5 * (vbar ** 2).sum(0)))
    expected = 0.5 * (wl[4] + wr[4] + sdl * sdr * cs * (wl[1] - wr[1]))
    flx, fp = O.riemann("roe", wl, wr, DIM1)
    assert np.allclose(fp, expected, rtol=1e-10, atol=1e-10)
    assert np.abs(flx).sum() > 0.0


def _euler_flux(w, ivx):
    rho, p = w[0], w[4]
    un = w[ivx]
    E = p / (GAMMA - 1) + 0.5 * rho * (w[1] ** 2 + w[2] ** 2 + w[3] ** 2)
    f = np.empty_like(w)
    f[0] = rho * un
    for v in (1, 2, 3):
        f[v] = rho * w[v] * un
    f[ivx] += p
    f[4] = (E + p) * un
    return f


@pytest.mark.parametrize("solver", ["lmars", "hllc", "roe"])
@pytest.mark.parametrize("dim", [3, 2, 1])
def test_consistency_equal_states(solver, dim):
    """F(w, w) is the physical Euler flux and the face pressure is p
    (the Roe case is tests/test_riemann.cpp:572-618 for the dry gas)."""
    wl, _ = _states(80 + dim)
    flx, fp = O.riemann(solver, wl, wl.copy(), dim)
    ivx = 4 - dim
    exact = _euler_flux(wl, ivx)
    assert np.allclose(flx, exact, rtol=1e-12, atol=1e-12)
    assert np.allclose(fp, wl[4], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("solver", ["lmars", "hllc", "roe"])
def test_mirror_symmetry(solver):
    """Reflecting x -> -x swaps the states, negates the normal velocity and
    negates every flux component except normal momentum."""
    wl, wr = _states(91)
    f, _ = O.riemann(solver, wl, wr, DIM1)
    ml, mr = wr.copy(), wl.copy()
    ml[1] *= -1
    mr[1] *= -1
    g, _ = O.riemann(solver, ml, mr, DIM1)
    sign = np.array([-1, 1, -1, -1, -1])[:, None, None, None]
    assert np.allclose(g, sign * f, rtol=1e-12, atol=1e-12)


# --- exact Riemann solver for an ideal gas (Toro, "Riemann Solvers and
# Numerical Methods for Fluid Dynamics", ch. 4) ---

def _exact_riemann(wl, wr, g, xi):
    dl, ul, pl = wl
    dr, ur, pr = wr
    cl, cr = np.sqrt(g * pl / dl), np.sqrt(g * pr / dr)

    def f(p, d, pk, ck):
        if p > pk:
            A, B = 2 / ((g + 1) * d), (g - 1) / (g + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

    lo, hi = 1e-8, 10 * max(pl, pr)
    for _ in range(200):
        pm = 0.5 * (lo + hi)
        if f(pm, dl, pl, cl) + f(pm, dr, pr, cr) + ur - ul > 0:
            hi = pm
        else:
            lo = pm
    um = 0.5 * (ul + ur) + 0.5 * (f(pm, dr, pr, cr) - f(pm, dl, pl, cl))
    out = np.empty((3, xi.size))
    for n, s in enumerate(xi):
        if s < um:  # left of contact
            if pm > pl:
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * pm / pl + (g - 1) / (2 * g))
                if s < sl:
                    out[:, n] = wl
                else:
                    d = dl * ((pm / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * pm / pl + 1))
                    out[:, n] = d, um, pm
            else:
                shl = ul - cl
                cml = cl * (pm / pl) ** ((g - 1) / (2 * g))
                stl = um - cml
                if s < shl:
                    out[:, n] = wl
                elif s > stl:
                    out[:, n] = dl * (pm / pl) ** (1 / g), um, pm
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    d = dl * (c / cl) ** (2 / (g - 1))
                    out[:, n] = d, u, pl * (c / cl) ** (2 * g / (g - 1))
        else:
            if pm > pr:
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * pm / pr + (g - 1) / (2 * g))
                if s > sr:
                    out[:, n] = wr
                else:
                    d = dr * ((pm / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * pm / pr + 1))
                    out[:, n] = d, um, pm
            else:
                shr = ur + cr
                cmr = cr * (pm / pr) ** ((g - 1) / (2 * g))
                str_ = um + cmr
                if s > shr:
                    out[:, n] = wr
                elif s < str_:
                    out[:, n] = dr * (pm / pr) ** (1 / g), um, pm
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    d = dr * (c / cr) ** (2 / (g - 1))
                    out[:, n] = d, u, pr * (c / cr) ** (2 * g / (g - 1))
    return out


def run_sod(tend=0.2, nx=256, cfl=0.9):
    c = O.sod_config(nx)
    w, u = O.initialize(c, O.sod_init(c))
    t = 0.0
    while t < tend - 1e-14:
        w = O.cons2prim(c, u)
        dt = min(O.max_time_step(c, w, cfl), tend - t)
        u, w = O.rk3_step(c, dt, u)
        t += dt
    return c, O.cons2prim(c, u)


def test_sod_against_exact_solution():
    """WENO5 + HLLC + SSP-RK3 on the Sod tube converges to the exact solution
    (density L1 error at t = 0.2 with 256 cells well below 1 %)."""
    c, w = run_sod()
    x = (np.arange(c.nx1) + 0.5) / c.nx1
    ex = _exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), GAMMA, (x - 0.5) / 0.2)
    rho = w[0, 0, 0, 3:-3]
    l1 = np.abs(rho - ex[0]).mean()
    assert l1 < 4e-3, l1
    # halving the resolution roughly doubles the (shock-dominated) error
    c2, w2 = run_sod(nx=128)
    x2 = (np.arange(128) + 0.5) / 128
    ex2 = _exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), GAMMA, (x2 - 0.5) / 0.2)
    l1_2 = np.abs(w2[0, 0, 0, 3:-3] - ex2[0]).mean()
    assert 1.4 < l1_2 / l1 < 3.0
