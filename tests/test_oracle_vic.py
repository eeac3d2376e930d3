"""Pins the oracle's vertically implicit correction (VIC).

The block Thomas solve is checked against a dense solve of the assembled
block-tridiagonal system; the MS-VIC redistribution replicates the reference's
own test (tests/test_backward_substitution.cpp:25-153); the rest are
invariants (column mass, limits).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O


def _straka_state(scheme, nx2=16, nx1=24, seed=3):
    c = O.straka_config(nx2=nx2, nx1=nx1, implicit_scheme=0)
    c.x2max = 25.6e3 * nx2 / 512
    w, u = O.initialize(c, O.straka_init(c))
    rng = np.random.default_rng(seed)
    # make the state less symmetric so every block entry is exercised
    w[1:4] += rng.uniform(-2.0, 2.0, size=w[1:4].shape)
    w, u = O.initialize(c, w)
    du, w2, _ = O.hydro_forward(c, 0.5, u)
    c.implicit_scheme = scheme
    return c, du, w2


def _dense(a, b, cc):
    nl, N, _ = a.shape
    M = np.zeros((nl * N, nl * N))
    for i in range(nl):
        M[i * N:(i + 1) * N, i * N:(i + 1) * N] = a[i]
        if i > 0:
            M[i * N:(i + 1) * N, (i - 1) * N:i * N] = b[i]
        if i < nl - 1:
            M[i * N:(i + 1) * N, (i + 1) * N:(i + 2) * N] = cc[i]
    return M


@pytest.mark.parametrize("scheme", [1, 9])
def test_block_thomas_matches_dense_solve(scheme):
    dt = 0.5
    c, du, w = _straka_state(scheme)
    out = O.implicit_correction(c, dt, du, w)
    g = c.nghost
    n3, n2, n1 = O.shape(c)
    N = 5 if scheme == 9 else 3
    vars_ = [0, 1, 4] if N == 3 else [0, 1, 2, 3, 4]
    dx1, dx2, dx3 = O.cartesian_widths(c)
    for j in (g, g + 5, n2 - g - 1):
        a, b, cc = O.vic_column_matrices(c, dt, w, 0, j)
        M = _dense(a, b, cc)
        rhs = np.stack([du[v, 0, j, g:-g] for v in vars_], axis=1).reshape(-1) / dt
        delta = np.linalg.solve(M, rhs).reshape(-1, N)
        tol = 1e-11  # observed ~1e-13 at cond(M) ~ 1e6
        for q, v in enumerate(vars_[1:], start=1):
            got = out[v, 0, j, g:-g]
            assert np.allclose(got, delta[:, q], rtol=tol, atol=tol * np.abs(delta[:, q]).max())
        # density: MS-VIC redistribution of the implicit mass tendency
        vol = dx3[0] * dx2[j] * dx1[g:-g]
        rho = w[0, 0, j, g:-g]
        ai = rho * vol
        b_dry = ((du[0, 0, j, g:-g] - delta[:, 0]) * vol).sum()
        exp = delta[:, 0] + rho * ai * b_dry / (ai * ai).sum()
        scale = np.abs(du[0, 0, j, g:-g]).max() + np.abs(delta[:, 0]).max()
        assert np.allclose(out[0, 0, j, g:-g], exp, rtol=0, atol=tol * scale)


@pytest.mark.parametrize("scheme", [1, 9])
def test_column_mass_is_conserved(scheme):
    dt = 0.5
    c, du, w = _straka_state(scheme)
    out = O.implicit_correction(c, dt, du, w)
    g = c.nghost
    before = du[0, 0, g:-g, g:-g].sum(axis=-1)
    after = out[0, 0, g:-g, g:-g].sum(axis=-1)
    scale = np.abs(du[0]).max() * c.nx1
    assert np.abs(after - before).max() < 1e-12 * scale


def test_small_dt_limit_leaves_du_unchanged():
    """As dt -> 0 the blocks are dominated by I/dt and the correction vanishes
    (delta -> du)."""
    c, du, w = _straka_state(1)
    du_small = du * 1e-9 / 0.5
    out = O.implicit_correction(c, 1e-9, du_small, w)
    g = c.nghost
    I = (slice(None), 0, slice(g, -g), slice(g, -g))
    assert np.allclose(out[I], du_small[I], rtol=1e-6, atol=1e-6 * np.abs(du_small).max())


def test_inverse3_vs_numpy():
    rng = np.random.default_rng(0)
    for _ in range(100):
        m = rng.normal(size=(3, 3)) + 3 * np.eye(3)
        assert np.allclose(O.inverse3(m), np.linalg.inv(m), rtol=1e-13, atol=1e-13)


def test_reference_redistribution_case():
    """tests/test_backward_substitution.cpp:25-153 (3 species, 4 layers,
    a = 0 so the back-substitution is the identity)."""
    L = O.lib()
    P = C.POINTER(C.c_double)
    nlayer, ny = 4, 3
    nhydro = 5 + ny
    stride1, stride2 = nlayer, 1
    du = np.zeros(nhydro * nlayer)
    w = np.zeros(nhydro * nlayer)
    vol = np.array([1.0, 2.0, 1.5, 0.5])
    corr = np.array([0.4, -0.2, 0.1, -0.3])
    rho = np.array([3.0, 2.5, 4.0, 1.8])
    a = np.zeros((nlayer, 9))
    delta = np.zeros((nlayer, 3))
    mass_fix = np.zeros(nhydro * nlayer)
    scratch = np.zeros(2 + ny)
    for i in range(nlayer):
        vapor, cloud, tracer = 0.10 + 0.02 * i, 0.05 + 0.01 * i, 0.02 + 0.005 * i
        dry = 1.0 - vapor - cloud - tracer
        et = 1.5 + 0.2 * i
        du[0 * stride1 + i] = et * dry
        du[1 * stride1 + i] = 10.0 + i
        du[4 * stride1 + i] = 20.0 + i
        du[5 * stride1 + i] = et * vapor
        du[6 * stride1 + i] = et * cloud
        du[7 * stride1 + i] = et * tracer
        w[5 * stride1 + i] = vapor
        w[6 * stride1 + i] = cloud
        w[7 * stride1 + i] = tracer
        w[0 * stride1 + i] = rho[i]
        delta[i] = (et + corr[i], 100.0 + i, 200.0 + i)
    original = du.copy()
    implicit_total = delta[:, 0].copy()
    sum_a2 = ((rho * vol) ** 2).sum()
    pp = lambda x: x.ctypes.data_as(P)
    L.oc_vic_backward_reduce(3, pp(du), pp(w), pp(a), pp(delta), pp(vol), pp(scratch),
                             0, nlayer - 1, 0, ny, stride1, stride2)
    for i in range(nlayer):
        L.oc_vic_redistribute_cell(3, pp(du), pp(w), pp(mass_fix), pp(delta), pp(vol),
                                   pp(scratch), i, 0, ny, stride1, stride2)

    def colint(u, var):
        return sum(vol[i] * u[var * stride1 + i] for i in range(nlayer))

    for var in (0, 5):
        assert abs(colint(du, var) - colint(original, var)) < 1e-1
    for var in (0, 5):
        b_j = 0.0
        for i in range(nlayer):
            frac = (1.0 - w[5 * stride1 + i] - w[6 * stride1 + i] - w[7 * stride1 + i]
                    if var == 0 else w[var * stride1 + i])
            b_j += (original[i] + original[5 * stride1 + i] + original[6 * stride1 + i]
                    - implicit_total[i]) * frac * vol[i]
        for i in range(nlayer):
            frac = (1.0 - w[5 * stride1 + i] - w[6 * stride1 + i]
                    if var == 0 else w[var * stride1 + i])
            et = sum(original[v * stride1 + i] for v in (0, 5, 6, 7))
            ai = rho[i] * vol[i]
            exp = (original[var * stride1 + i] + (implicit_total[i] - et) * frac
                   + rho[i] * ai * b_j / sum_a2)
            assert abs(du[var * stride1 + i] - exp) < 1e-1
    for i in range(nlayer):
        et = sum(original[v * stride1 + i] for v in (0, 5, 6, 7))
        assert abs(mass_fix[i] - (implicit_total[i] - et)) < 1e-12
        for var in (6, 7):
            exp = original[var * stride1 + i] + (implicit_total[i] - et) * w[var * stride1 + i]
            assert abs(du[var * stride1 + i] - exp) < 1e-12
    # momentum / energy take the implicit solution
    assert np.allclose(du[1 * stride1:2 * stride1], delta[:, 1])
    assert np.allclose(du[4 * stride1:5 * stride1], delta[:, 2])
