"""Pins the oracle's EOS, divergence, gravity revisions, boundary functions,
SSP-RK3 driver and the composed Hydro RHS by closed forms and invariants."""
import numpy as np
import pytest

import oracle as O

R, CP = 287.0, 1004.0


def _rand_prim(shape, seed=0):
    rng = np.random.default_rng(seed)
    w = np.empty((5,) + shape)
    w[0] = rng.uniform(0.5, 2.0, shape)
    w[1:4] = rng.uniform(-3.0, 3.0, (3,) + shape)
    w[4] = rng.uniform(0.5e5, 1.5e5, shape)
    return w


def test_cons_prim_round_trip_and_closed_form():
    # W->U->W round trip (tests/test_eos.cpp:58-61 pattern) and the ideal-gas
    # closed forms (src/eos/ideal_gas_impl.h:15-35, src/eos/ideal_gas.cpp:67-88)
    c = O.make_config(nx1=4, nx2=3, nx3=2, nghost=1, gammad=1.4)
    w = _rand_prim((4, 5, 6))
    u = O.prim2cons(c, w)
    assert np.allclose(u[0], w[0], rtol=0, atol=0)
    assert np.allclose(u[1:4], w[1:4] * w[0], rtol=1e-15)
    e = w[4] / 0.4 + 0.5 * w[0] * (w[1:4] ** 2).sum(0)
    assert np.allclose(u[4], e, rtol=1e-14)
    w2 = O.cons2prim(c, u)
    assert np.allclose(w2, w, rtol=1e-12, atol=1e-12)


def test_divergence_uniform_and_telescoping():
    c = O.make_config(nx1=8, nx2=6, nx3=5, nghost=3, x1max=2.0, x2max=3.0, x3max=0.5)
    n3, n2, n1 = O.shape(c)
    f = [np.full((5, n3, n2, n1), 3.7) for _ in range(3)]
    d = O.divergence(c, *f)
    assert np.abs(d).max() < 1e-12
    rng = np.random.default_rng(2)
    f = [rng.normal(size=(5, n3, n2, n1)) for _ in range(3)]
    d = O.divergence(c, *f)
    d1, d2, d3 = O.cartesian_widths(c)
    g = c.nghost
    vol = d1[0] * d2[0] * d3[0]
    total = (d[O.interior(c)] * vol).sum(axis=(1, 2, 3))
    a1, a2, a3 = d2[0] * d3[0], d1[0] * d3[0], d1[0] * d2[0]
    I = slice(g, -g)
    bnd = (a1 * (f[0][:, I, I, n1 - g] - f[0][:, I, I, g]).sum(axis=(1, 2))
           + a2 * (f[1][:, I, n2 - g, I] - f[1][:, I, g, I]).sum(axis=(1, 2))
           + a3 * (f[2][:, n3 - g, I, I] - f[2][:, g, I, I]).sum(axis=(1, 2)))
    assert np.allclose(total, bnd, rtol=1e-12, atol=1e-12)


def _isentropic_column(nx1=32, dz=100.0, gamma=CP / (CP - R)):
    g = 9.81
    theta, p0 = 300.0, 1e5
    ng = 3
    z = (np.arange(nx1 + 2 * ng) - ng + 0.5) * dz
    T = theta - g * z / CP
    p = p0 * (T / theta) ** (CP / R)
    w = np.zeros((5, 1, 1, nx1 + 2 * ng))
    w[0, 0, 0] = p / (R * T)
    w[4, 0, 0] = p
    return w, gamma, g, dz


def test_gravity_ghost_refill_is_isentropic_closed_form():
    """src/hydro/hydro.cpp:231-269: for an isentropic atmosphere
    p^((g-1)/g) is linear in z, so the refilled ghosts equal the analytic
    column (cp = R g/(g-1) makes the discrete profile exactly isentropic)."""
    w, gamma, g, dz = _isentropic_column()
    exact = w.copy()
    w[:, :, :, :3] = 0.0
    w[:, :, :, -3:] = 0.0
    import ctypes as C
    L = O.lib()
    P = C.POINTER(C.c_double)
    dx1v = np.full(w.shape[-1], dz)
    w = np.ascontiguousarray(w)
    L.oc_revise_x1_ghost(w.ctypes.data_as(P), 5, 1, 1, w.shape[-1], 3, -g, gamma,
                         dx1v.ctypes.data_as(P))
    assert np.allclose(w[0], exact[0], rtol=1e-10)
    assert np.allclose(w[4], exact[4], rtol=1e-10)


def test_hydrostatic_column_stays_at_rest():
    """Cartesian analogue of tests/test_hydrostatic.cpp:128-167: an
    isentropic column with gravity stays at rest, max |v1| < 1e-8 after 100
    RK3 steps (explicit and with the implicit vertical solve)."""
    for scheme, dt in ((0, 0.2), (1, 1.0), (9, 1.0)):
        c = O.make_config(nx1=64, nghost=3, x1min=0.0, x1max=6400.0,
                          gammad=CP / (CP - R), recon_vertical=("weno5", False, False),
                          riemann="lmars", grav1=-9.81, implicit_scheme=scheme)
        w, *_ = _isentropic_column(64)
        w, u = O.initialize(c, w)
        for _ in range(100):
            u, w = O.rk3_step(c, dt, u)
        w = O.cons2prim(c, u)
        assert np.abs(w[1, 0, 0, 3:-3]).max() < 1e-8


@pytest.mark.parametrize("scheme", [0, 1])
def test_straka_mass_conserved_to_roundoff(scheme):
    c = O.straka_config(nx2=64, nx1=32, implicit_scheme=scheme)
    c.x2max = 25.6e3 * 64 / 512 * 4
    w, u = O.initialize(c, O.straka_init(c))
    I = O.interior(c)
    m0 = u[0][I[1:]].sum()
    for _ in range(5):
        u, w = O.rk3_step(c, 0.5, u)
    m1 = u[0][I[1:]].sum()
    assert abs(m1 - m0) / m0 < 1e-13


def test_boundary_functions():
    """src/bc/bc_func.cpp:7-63 on a random block."""
    c = O.make_config(nx1=5, nx2=6, nghost=3,
                      bc=("reflecting", "outflow", "periodic", "periodic", None, None))
    n3, n2, n1 = O.shape(c)
    rng = np.random.default_rng(1)
    v = rng.normal(size=(5, n3, n2, n1))
    out = O.apply_bc(c, v, O.CONSERVED)
    g = 3
    # x1 inner reflecting: mirror, IVX negated (applied first, then x2 periodic)
    j = slice(g, n2 - g)
    for q in range(g):
        exp = v[:, :, j, 2 * g - 1 - q].copy()
        exp[1] *= -1
        assert np.array_equal(out[:, :, j, q], exp)
        assert np.array_equal(out[:, :, j, n1 - g + q], v[:, :, j, n1 - g - 1])
    # x2 periodic over every x1 (including x1 ghosts)
    for q in range(g):
        assert np.array_equal(out[:, :, q, :], out[:, :, n2 - 2 * g + q, :])
        assert np.array_equal(out[:, :, n2 - g + q, :], out[:, :, g + q, :])


def _advect(nx, dt, tend, cfl_dt=None):
    c = O.make_config(nx1=nx, nghost=3, gammad=1.4, recon_vertical=("weno5", False, True),
                      riemann="hllc", bc=("periodic",) * 6)
    n3, n2, n1 = O.shape(c)
    x = (np.arange(n1) - 3 + 0.5) / nx
    w = np.zeros((5, 1, 1, n1))
    w[0, 0, 0] = 1.0 + 0.2 * np.sin(2 * np.pi * x)
    w[1] = 1.0
    w[4] = 1.0
    w, u = O.initialize(c, w)
    nsteps = int(round(tend / dt))
    for _ in range(nsteps):
        u, w = O.rk3_step(c, dt, u)
    return u


def test_rk3_third_order_in_time():
    """SSP-RK3 (harp integrator weights, not vendored): with the spatial
    discretisation fixed, the error against a tiny-dt solution falls by ~8
    when dt halves."""
    nx, tend = 32, 0.25
    ref = _advect(nx, tend / 800, tend)
    e = []
    for n in (20, 40):
        u = _advect(nx, tend / n, tend)
        e.append(np.abs(u[0] - ref[0]).max())
    order = np.log2(e[0] / e[1])
    assert 2.6 < order < 3.4, (e, order)


def test_max_time_step_closed_form():
    c = O.sod_config(64)
    w = O.sod_init(c)
    cs = np.sqrt(1.4 * w[4] / w[0])
    exp = 0.9 * (1 / 64 / (np.abs(w[1]) + cs))[0, 0, 3:-3].min()
    assert abs(O.max_time_step(c, w, 0.9) - exp) < 1e-15


def test_custom_boundaries_leave_the_callers_ghosts():
    """custom_inner / custom_outer (src/bc/bc_func.cpp:4-5) are no-ops: the
    x2 ghosts keep the caller's values while the x1 walls reflect."""
    c = O.make_config(nx1=8, nx2=6, nghost=3,
                      bc=("reflecting", "reflecting", "custom", "custom", "periodic", "periodic"))
    rng = np.random.default_rng(3)
    var = rng.standard_normal((5,) + O.shape(c))
    out = O.apply_bc(c, var, O.CONSERVED)
    g = 3
    assert np.array_equal(out[:, :, :g, g:-g], var[:, :, :g, g:-g])
    assert np.array_equal(out[:, :, -g:, g:-g], var[:, :, -g:, g:-g])
    assert np.array_equal(out[1, :, g:-g, g - 1], -var[1, :, g:-g, g])
