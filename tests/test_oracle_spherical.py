"""Pins the oracle's spherical-polar geometry and coordinate sources
(src/coord/spherical_polar.cpp).

Geometry is checked against the integrals that define it (scipy quadrature
and closed shell volumes), the flux divergence and the source coefficients
against Gauss's theorem and the conservation laws the reference's source
terms are built to keep (mass, and angular momentum about the origin and
about the polar axis), and the coordinate forward against the reference's own
tests (tests/test_coordinate.cpp:591-669).  A state at rest in uniform
pressure must have no tendency at all."""
import numpy as np
import pytest
from scipy.integrate import quad

import oracle as O

BC = ("reflecting", "reflecting", "reflecting", "reflecting", "periodic", "periodic")


def _sph(nx1=10, nx2=8, nx3=6, **kw):
    opts = dict(nx1=nx1, nx2=nx2, nx3=nx3, nghost=3, x1min=1.0, x1max=2.0,
                x2min=0.25 * np.pi, x2max=0.75 * np.pi, x3min=-0.5 * np.pi, x3max=0.5 * np.pi,
                coord="spherical-polar", riemann="hllc", bc=BC)
    opts.update(kw)
    return O.make_config(**opts)


def _faces(lo, hi, nx, ng=3):
    d = (hi - lo) / nx
    return np.linspace(lo - ng * d, hi + ng * d, nx + 2 * ng + 1)


def test_centres_are_volume_weighted():
    """x1v = int r r^2 dr / int r^2 dr, x2v = int t sin t dt / int sin t dt
    (src/coord/spherical_polar.cpp:17-28)."""
    c = _sph()
    r, t = _faces(1.0, 2.0, 10), _faces(0.25 * np.pi, 0.75 * np.pi, 8)
    x1v, x2v = O.geometry(c, "x1v"), O.geometry(c, "x2v")
    for i in range(len(r) - 1):
        num = quad(lambda x: x ** 3, r[i], r[i + 1])[0]
        den = quad(lambda x: x ** 2, r[i], r[i + 1])[0]
        assert abs(x1v[i] - num / den) < 1e-12 * abs(x1v[i])
    for j in range(len(t) - 1):
        num = quad(lambda x: x * np.sin(x), t[j], t[j + 1])[0]
        den = quad(np.sin, t[j], t[j + 1])[0]
        assert abs(x2v[j] - num / den) < 1e-12
    # center_distances: dx1v[0] = dx1f[0], then differences of x1v
    dx1v = O.geometry(c, "dx1v")
    assert dx1v[0] == pytest.approx(r[1] - r[0], rel=1e-13)
    assert np.allclose(dx1v[1:], np.diff(x1v), rtol=1e-13)


def test_volumes_and_areas_integrate_the_sector():
    c = _sph()
    r = _faces(1.0, 2.0, 10)
    t = _faces(0.25 * np.pi, 0.75 * np.pi, 8)
    p = _faces(-0.5 * np.pi, 0.5 * np.pi, 6)
    vol = O.geometry(c, "vol")
    shell = (r[-1] ** 3 - r[0] ** 3) / 3 * abs(np.cos(t[0]) - np.cos(t[-1])) * (p[-1] - p[0])
    assert vol.sum() == pytest.approx(shell, rel=1e-13)
    a1, a2, a3 = O.geometry(c, "area1"), O.geometry(c, "area2"), O.geometry(c, "area3")
    solid = abs(np.cos(t[0]) - np.cos(t[-1])) * (p[-1] - p[0])
    assert np.allclose(a1.sum(axis=(0, 1)), r ** 2 * solid, rtol=1e-13)
    ring = 0.5 * (r[-1] ** 2 - r[0] ** 2) * (p[-1] - p[0])
    assert np.allclose(a2.sum(axis=(0, 2)), ring * np.abs(np.sin(t)), rtol=1e-13)
    wedge = 0.5 * (r[-1] ** 2 - r[0] ** 2) * (t[-1] - t[0])
    assert np.allclose(a3.sum(axis=(1, 2)), wedge, rtol=1e-13)


def _rest_div(c, pre=2.5, rho=1.3):
    n3, n2, n1 = O.shape(c)
    w = np.zeros((5, n3, n2, n1))
    w[0], w[4] = rho, pre
    f = np.zeros_like(w)
    f[1] = pre  # momentum fluxes of a state at rest: the pressure
    f2 = np.zeros_like(w)
    f2[2] = pre
    f3 = np.zeros_like(w)
    f3[3] = pre
    return w, f, f2, f3


def test_gauss_theorem_fixes_the_pressure_sources():
    """Uniform pressure at rest: the pressure flux through a cell's faces
    equals the geometric source (div e_r = 2 / r, div e_theta = cot(theta) / r
    integrated over the cell), so the momentum tendency vanishes, with the
    2 p / r source and with the face pressure alike."""
    c = _sph()
    w, f1, f2, f3 = _rest_div(c)
    I = O.interior(c)[1:]
    for fp1 in (None, np.full(w.shape[1:], 2.5)):
        div = O.coord_forward(c, w, f1, f2, f3, fp1)
        for v in range(5):
            assert np.abs(div[v][I]).max() < 1e-12 * 2.5, v


def test_reference_face_pressure_cases():
    """tests/test_coordinate.cpp:591-669: with the face pressure given, the
    x1 momentum source turns the area-weighted pressure flux into the plain
    face-pressure gradient, and the cell pressure no longer enters it."""
    c = O.make_config(nx1=4, nx2=4, nx3=4, nghost=1, x1min=1.0, x1max=2.0,
                      x2min=0.25 * np.pi, x2max=0.75 * np.pi, x3min=-0.5 * np.pi,
                      x3max=0.5 * np.pi, coord="spherical-polar", recon_vertical=("plm", 0, 0),
                      bc=BC)
    n3, n2, n1 = O.shape(c)
    x1f = _faces(1.0, 2.0, 4, 1)
    prim = np.zeros((5, n3, n2, n1))
    prim[0] = 1.0
    fp = np.broadcast_to(x1f[:n1], (n3, n2, n1)).copy()
    f1 = np.zeros_like(prim)
    f1[1] = fp
    zero = np.zeros_like(prim)
    div = O.coord_forward(c, prim, f1, zero, zero, fp)
    si, ei = 1, n1 - 1
    dx1f = np.diff(x1f)
    expect = (fp[..., si + 1:ei + 1] - fp[..., si:ei]) / dx1f[si:ei]
    assert np.allclose(div[1][..., si:ei], expect, rtol=1e-12, atol=1e-12)

    lo, hi = prim.copy(), prim.copy()
    lo[4], hi[4] = 3.0, 11.0
    face = np.full((n3, n2, n1), 5.0)
    d_lo, d_hi = O.coord_forward(c, lo, zero, zero, zero), O.coord_forward(c, hi, zero, zero, zero)
    assert not np.allclose(d_lo[1], d_hi[1], rtol=1e-6, atol=1e-6)
    d_lo = O.coord_forward(c, lo, zero, zero, zero, face)
    d_hi = O.coord_forward(c, hi, zero, zero, zero, face)
    assert np.allclose(d_lo[1], d_hi[1], rtol=1e-12, atol=1e-12)


def test_mass_divergence_telescopes():
    c = _sph()
    rng = np.random.default_rng(5)
    n3, n2, n1 = O.shape(c)
    f1, f2, f3 = (rng.standard_normal((5, n3, n2, n1)) for _ in range(3))
    w = np.ones((5, n3, n2, n1))
    div = O.coord_forward(c, w, f1, f2, f3)
    vol = O.geometry(c, "vol")
    a1, a2, a3 = O.geometry(c, "area1"), O.geometry(c, "area2"), O.geometry(c, "area3")
    g = 3
    s = slice(g, -g)
    lhs = (vol[s, s, s] * div[0][s, s, s]).sum()
    rhs = ((a1[s, s, n1 - g] * f1[0][s, s, n1 - g]).sum() - (a1[s, s, g] * f1[0][s, s, g]).sum()
           + (a2[s, n2 - g, s] * f2[0][s, n2 - g, s]).sum() - (a2[s, g, s] * f2[0][s, g, s]).sum()
           + (a3[n3 - g, s, s] * f3[0][n3 - g, s, s]).sum() - (a3[g, s, s] * f3[0][g, s, s]).sum())
    assert lhs == pytest.approx(rhs, rel=1e-12, abs=1e-12 * np.abs(vol * div[0]).sum())


def test_radial_flux_sources_conserve_angular_momentum():
    """The x1-flux sources of the transverse momenta turn the radial flux
    difference into (rp^3 F+ - rm^3 F-) / ((rm + rp)/2 Vr), the form that
    conserves r x (rho v).  The reference scales them with the face area of
    the whole cell, so the identity holds where that solid angle is 1: an
    equatorial band (sin theta equal on both faces) with dphi = 1 / |dcos|."""
    a = 0.3
    c = _sph(nx1=12, nx2=1, nx3=1, x2min=0.5 * np.pi - a, x2max=0.5 * np.pi + a,
             x3min=0.0, x3max=1.0 / (2 * np.sin(a)))
    n3, n2, n1 = O.shape(c)
    rng = np.random.default_rng(8)
    f1 = rng.standard_normal((5, n3, n2, n1))
    w = np.zeros((5, n3, n2, n1))
    w[0] = 1.0 + rng.random((n3, n2, n1))
    w[2:4] = rng.standard_normal((2, n3, n2, n1))
    w[4] = 1.0
    div = O.coord_forward(c, w, f1, None, None)
    r = _faces(1.0, 2.0, 12)
    g = 3
    vr = (r[1:] ** 3 - r[:-1] ** 3) / 3
    weight = 0.5 * (r[:-1] + r[1:]) * vr
    for v in (2, 3):
        lhs = (weight[g:-g] * div[v][0, 0, g:-g]).sum()
        rhs = r[n1 - g] ** 3 * f1[v][0, 0, n1 - g] - r[g] ** 3 * f1[v][0, 0, g]
        scale = np.abs(weight * div[v][0, 0]).sum()
        assert lhs == pytest.approx(rhs, abs=1e-12 * scale), v


def test_polar_flux_source_conserves_axial_angular_momentum():
    """The x2-flux source of the azimuthal momentum turns the polar flux
    difference into 2 (sin+^2 F+ - sin-^2 F-) / ((sin- + sin+) |dcos|) (R/Vr),
    the form that conserves r sin(theta) rho v_phi, where the reference's
    extra factor 0.5 (rp^2 - rm^2) dphi is 1: checked on the one radial row
    where dphi makes it so."""
    c0 = _sph(nx1=4, nx2=10, nx3=1)
    r = _faces(1.0, 2.0, 4)
    i0 = 4
    R = 0.5 * (r[i0 + 1] ** 2 - r[i0] ** 2)
    c = _sph(nx1=4, nx2=10, nx3=1, x3min=0.0, x3max=1.0 / R)
    del c0
    n3, n2, n1 = O.shape(c)
    rng = np.random.default_rng(9)
    f2 = np.zeros((5, n3, n2, n1))
    f2[3] = rng.standard_normal((n3, n2, n1))
    zero = np.zeros_like(f2)
    w = np.zeros_like(f2)
    w[0], w[4] = 1.0, 1.0
    div = O.coord_forward(c, w, zero, f2, None)
    t = _faces(0.25 * np.pi, 0.75 * np.pi, 10)
    g = 3
    sm, sp = np.abs(np.sin(t[:-1])), np.abs(np.sin(t[1:]))
    dcos = np.abs(np.cos(t[:-1]) - np.cos(t[1:]))
    vr = (r[i0 + 1] ** 3 - r[i0] ** 3) / 3
    weight = (vr / R) * 0.5 * (sm + sp) * dcos
    # the polar-momentum part of the tendency of a state at rest at unit
    # pressure is removed with it (it is the -src1_i src1_j p term only)
    lhs = (weight[g:-g] * div[3][0, g:-g, i0]).sum()
    rhs = np.sin(t[n2 - g]) ** 2 * f2[3][0, n2 - g, i0] - np.sin(t[g]) ** 2 * f2[3][0, g, i0]
    scale = np.abs(weight * div[3][0, :, i0]).sum()
    assert lhs == pytest.approx(rhs, abs=1e-12 * scale)


@pytest.mark.parametrize("solver", ["lmars", "hllc", "roe"])
def test_rest_state_has_no_tendency(solver):
    """hydro_forward of a uniform state at rest: every interior tendency
    vanishes (reflecting x1/x2 walls, periodic x3)."""
    c = _sph(riemann=solver)
    n3, n2, n1 = O.shape(c)
    w = np.zeros((5, n3, n2, n1))
    w[0], w[4] = 1.3, 2.5
    u = O.prim2cons(c, w)
    du, _, _ = O.hydro_forward(c, 1e-3, u)
    assert np.abs(du[O.interior(c)]).max() < 1e-14


@pytest.mark.parametrize("scheme", [1, 9])
def test_implicit_correction_conserves_column_mass(scheme):
    """MS-VIC redistribution on r-dependent volumes: sum_i V_i du_i[rho] of
    every column is what the explicit tendency had."""
    c = _sph(nx1=12, nx2=6, nx3=4, grav1=-0.5)
    n3, n2, n1 = O.shape(c)
    rng = np.random.default_rng(11)
    w = np.zeros((5, n3, n2, n1))
    w[0] = 1.0 + 0.1 * rng.random((n3, n2, n1))
    w[1:4] = 0.05 * rng.standard_normal((3, n3, n2, n1))
    w[4] = 1.0 + 0.1 * rng.random((n3, n2, n1))
    w, u = O.initialize(c, w)
    du, w2, _ = O.hydro_forward(c, 0.05, u)
    c.implicit_scheme = scheme
    out = O.implicit_correction(c, 0.05, du, w2)
    vol = O.geometry(c, "vol")
    g = 3
    s = slice(g, -g)
    before = (vol[s, s, s] * du[0][s, s, s]).sum(axis=-1)
    after = (vol[s, s, s] * out[0][s, s, s]).sum(axis=-1)
    scale = (vol[s, s, s] * np.abs(du[0][s, s, s])).sum(axis=-1).max()
    assert np.abs(after - before).max() < 1e-12 * scale
    assert not np.allclose(out[1], du[1])


def test_rk3_conserves_mass_in_a_closed_sector():
    """Reflecting r and theta walls, periodic phi: total mass sum_V rho V is
    constant to roundoff over RK3 steps with the implicit vertical solve."""
    c = _sph(nx1=12, nx2=6, nx3=4, grav1=-0.5, implicit_scheme=1)
    n3, n2, n1 = O.shape(c)
    rng = np.random.default_rng(12)
    w = np.zeros((5, n3, n2, n1))
    w[0] = 1.0 + 0.1 * rng.random((n3, n2, n1))
    w[1:4] = 0.05 * rng.standard_normal((3, n3, n2, n1))
    w[4] = 1.0 + 0.1 * rng.random((n3, n2, n1))
    w, u = O.initialize(c, w)
    vol = O.geometry(c, "vol")
    I = O.interior(c)[1:]
    m0 = (vol[I] * u[0][I]).sum()
    for _ in range(5):
        u, w = O.rk3_step(c, 0.02, u)
    assert (vol[I] * u[0][I]).sum() == pytest.approx(m0, rel=1e-13)
