"""CPU: the oracle's gnomonic-equiangle panel (src/coord/gnomonic_equiangle.cpp)
pinned by the geometry of the sphere, independently of the formulas' text.

A panel point (alpha, beta) is P = (1, tan alpha, tan beta) / |.| on the unit
sphere; lines of constant alpha or beta are great circles.  Checks:

* the CFL widths are r times the great-circle arcs between a cell's faces;
* the cosines of the cell and face angles are those between the unit
  tangents dP/dalpha and dP/dbeta (complex-step derivatives);
* the x2 / x3 faces are planar annular sectors: their areas are exactly
  (r_+^2 - r_-^2) / 2 times the arc;
* the x1 face areas over a panel converge to the panel's solid angle 4 pi / 6
  times r^2, and the volumes to the shell sector, at second order;
* the ideal-gas EOS lowers velocities to covariant momenta with the cell's
  cosine and its energy holds the metric norm |v|^2 = v1^2 + v2^2 + v3^2 +
  2 cos v2 v3;
* the face frames give the velocity components along the face normal and
  tangent in 3-D, and the fluxes back the covariant components of the 3-D
  momentum flux;
* a uniform pressure at rest has no tendency with LMARS and HLLC (Gauss's
  theorem for the face pressures and the metric sources), and RK3 steps with
  reflecting walls conserve the mass.
Each planted mistake fails at least one of them."""
import math

import numpy as np
import pytest

import oracle as O

PI4 = math.pi / 4


def panel(nx1=4, nx=12, riemann="lmars", bc=("outflow",) * 6, r=(1.0, 1.5), lo=-PI4, hi=PI4,
          **kw):
    return O.make_config(nx1=nx1, nx2=nx, nx3=nx, nghost=3, x1min=r[0], x1max=r[1],
                         x2min=lo, x2max=hi, x3min=lo, x3max=hi, riemann=riemann,
                         coord="gnomonic-equiangle", bc=bc, **kw)


def faces(lo, hi, nx, ng=3):
    d = (hi - lo) / nx
    return np.linspace(lo - ng * d, hi + ng * d, nx + 2 * ng + 1)


def point(a, b):
    v = np.stack(np.broadcast_arrays(np.ones_like(np.tan(a) * np.tan(b)), np.tan(a), np.tan(b)))
    return v / np.sqrt((v * v).sum(0))


def arc(p, q):
    return np.arctan2(np.linalg.norm(np.cross(p, q, axis=0), axis=0), (p * q).sum(0))


def tangents(a, b, h=1e-30):
    """unit dP/dalpha, dP/dbeta by complex steps"""
    ta = point(a + 1j * h, b).imag / h
    tb = point(a, b + 1j * h).imag / h
    return ta / np.linalg.norm(ta, axis=0), tb / np.linalg.norm(tb, axis=0)


def axes(c):
    x1f = faces(c.x1min, c.x1max, c.nx1)
    x2f = faces(c.x2min, c.x2max, c.nx2)
    x3f = faces(c.x3min, c.x3max, c.nx3)
    return x1f, x2f, x3f, 0.5 * (x1f[1:] + x1f[:-1]), 0.5 * (x2f[1:] + x2f[:-1]), \
        0.5 * (x3f[1:] + x3f[:-1])


def test_cfl_widths_are_great_circle_arcs():
    c = panel()
    x1f, x2f, x3f, x1v, x2v, x3v = axes(c)
    B, A = np.meshgrid(x3v, x2f, indexing="ij")  # [n3][n2 + 1]
    a2 = arc(point(A[:, :-1], B[:, :-1]), point(A[:, 1:], B[:, 1:]))
    B, A = np.meshgrid(x3f, x2v, indexing="ij")  # [n3 + 1][n2]
    a3 = arc(point(A[:-1], B[:-1]), point(A[1:], B[1:]))
    w2, w3 = O.geometry(c, "width2"), O.geometry(c, "width3")
    assert np.allclose(w2, x1v * a2[..., None], rtol=1e-12, atol=0)
    assert np.allclose(w3, x1v * a3[..., None], rtol=1e-12, atol=0)
    assert np.allclose(O.geometry(c, "width1"), (c.x1max - c.x1min) / c.nx1, rtol=1e-15)


def test_cfl_width_check_fails_on_planted_mistake():
    c = panel()
    x1f, x2f, x3f, x1v, x2v, x3v = axes(c)
    B, A = np.meshgrid(x3f, x2v, indexing="ij")
    a3 = arc(point(A[:-1], B[:-1]), point(A[1:], B[1:]))
    with O.plant("gno_ang3"):
        w3 = O.geometry(c, "width3")
    # NaN (acos beyond 1) counts as a failure of the check too
    assert not (np.abs(w3 / (x1v * a3[..., None]) - 1) <= 1e-3).all()


def test_angles_between_coordinate_lines():
    c = panel()
    x1f, x2f, x3f, x1v, x2v, x3v = axes(c)
    for what, (aa, bb) in {"cell": (x2v, x3v), "face2": (x2f, x3v), "face3": (x2v, x3f)}.items():
        B, A = np.meshgrid(bb, aa, indexing="ij")
        ta, tb = tangents(A, B)
        cos = (ta * tb).sum(0)
        assert np.allclose(O.geometry(c, "cos_" + what), cos, rtol=0, atol=1e-14), what
        assert np.allclose(O.geometry(c, "sin_" + what), np.sqrt(1 - cos * cos), rtol=0,
                           atol=1e-14), what


def test_side_faces_are_planar_sectors():
    c = panel()
    x1f, x2f, x3f, x1v, x2v, x3v = axes(c)
    ring = 0.5 * (x1f[1:] ** 2 - x1f[:-1] ** 2)
    B, A = np.meshgrid(x3f, x2f, indexing="ij")  # corners [n3 + 1][n2 + 1]
    P = point(A, B)
    arc2 = arc(P[:, :-1, :], P[:, 1:, :])  # along x3 on each x2 line: [n3][n2 + 1]
    arc3 = arc(P[:, :, :-1], P[:, :, 1:])  # along x2 on each x3 line: [n3 + 1][n2]
    assert np.allclose(O.geometry(c, "area2"), arc2[..., None] * ring, rtol=1e-12, atol=0)
    assert np.allclose(O.geometry(c, "area3"), arc3[..., None] * ring, rtol=1e-12, atol=0)


def test_panel_solid_angle_and_volume_converge_at_second_order():
    errs_a, errs_v = [], []
    for n in (8, 16, 32):
        c = panel(nx1=n, nx=n)
        g = 3
        x1f = faces(c.x1min, c.x1max, c.nx1)
        a1 = O.geometry(c, "area1")[g:-g, g:-g, g]  # x1 face at r = 1
        errs_a.append(abs(a1.sum() / x1f[g] ** 2 / (4 * math.pi / 6) - 1))
        v = O.geometry(c, "vol")[g:-g, g:-g, g:-g].sum()
        errs_v.append(abs(v / (4 * math.pi / 6 * (1.5 ** 3 - 1.0) / 3) - 1))
    for e in (errs_a, errs_v):
        assert e[2] < 1e-3, e
        assert 3.5 < e[0] / e[1] < 4.5 and 3.5 < e[1] / e[2] < 4.5, e


def _random_prim(c, seed=0):
    rng = np.random.default_rng(seed)
    n3, n2, n1 = O.shape(c)
    w = np.empty((5, n3, n2, n1))
    w[0] = rng.uniform(0.8, 1.2, (n3, n2, n1))
    w[1:4] = rng.uniform(-0.3, 0.3, (3, n3, n2, n1))
    w[4] = rng.uniform(0.8, 1.2, (n3, n2, n1))
    return w


def eos_errors(c):
    w = _random_prim(c)
    n = w[0].size
    u = np.empty_like(w)
    O.lib().oc_prim2cons(O.C.byref(c), O._p(w.copy()), O._p(u), n)
    cos = O.geometry(c, "cos_cell")[..., None]
    rho, v1, v2, v3, p = w
    m2, m3 = rho * (v2 + cos * v3), rho * (v3 + cos * v2)
    e = p / (c.gammad - 1) + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3 + 2 * cos * v2 * v3)
    e_mom = max(np.abs(u[2] - m2).max(), np.abs(u[3] - m3).max())
    back = O.cons2prim(c, u)
    e_round = np.abs(back - w).max()
    return max(e_mom, np.abs(u[4] - e).max() / np.abs(e).max(), e_round)


def test_eos_lowers_velocities_and_keeps_the_metric_norm():
    assert eos_errors(panel()) < 1e-14


def test_eos_check_fails_on_planted_mistake():
    with O.plant("gno_no_raise"):
        assert eos_errors(panel()) > 1e-3


def _frame_basis(c, dim):
    """Unit normal and the tangent(s) of dim's faces in 3-D, with the unit
    tangents e_alpha, e_beta at the same points ([3][n3][n2] each)."""
    x1f, x2f, x3f, x1v, x2v, x3v = axes(c)
    if dim == 2:
        B, A = np.meshgrid(x3v, x2f[:-1], indexing="ij")
    elif dim == 1:
        B, A = np.meshgrid(x3f[:-1], x2v, indexing="ij")
    else:
        B, A = np.meshgrid(x3v, x2v, indexing="ij")
    ea, eb = tangents(A, B)
    rad = point(A, B)
    return ea, eb, rad


def frame_errors(c):
    w = _random_prim(c, seed=3)
    err = 0.0
    for dim in (3, 2, 1):
        ea, eb, rad = _frame_basis(c, dim)
        v2, v3 = w[2], w[3]
        vel = v2 * ea[..., None] + v3 * eb[..., None]  # [3][n3][n2][n1]
        loc = O.gno_prim2local(c, w, dim)
        if dim == 2:  # x2 face: normal in the tangent plane, perpendicular to e_beta
            n = np.cross(eb, rad, axis=0)
            n *= np.sign((n * ea).sum(0))
            t = eb
            got_n, got_t = loc[2], loc[3]
        elif dim == 1:  # x3 face: normal perpendicular to e_alpha
            n = np.cross(rad, ea, axis=0)
            n *= np.sign((n * eb).sum(0))
            t = ea
            got_n, got_t = loc[3], loc[2]
        else:  # x1 face: tangent plane spanned by e_alpha and its normal
            t = ea
            n = np.cross(rad, ea, axis=0)
            n *= np.sign((n * eb).sum(0))
            got_n, got_t = loc[3], loc[2]
        ref_n = (vel * n[..., None]).sum(0)
        ref_t = (vel * t[..., None]).sum(0)
        err = max(err, np.abs(got_n - ref_n).max(), np.abs(got_t - ref_t).max())
        # a face-frame flux (F_n along n, F_t along t) back to covariant
        # components M . e_alpha, M . e_beta of the 3-D momentum flux
        f = np.zeros_like(w)
        fn, ft = _random_prim(c, seed=7)[1:3]
        if dim == 2:
            f[2], f[3] = fn, ft
        elif dim == 1:
            f[3], f[2] = fn, ft
        else:
            f[2], f[3] = ft, fn
        M = fn * n[..., None] + ft * t[..., None]
        g = O.gno_flux2global(c, f, dim)
        err = max(err, np.abs(g[2] - (M * ea[..., None]).sum(0)).max(),
                  np.abs(g[3] - (M * eb[..., None]).sum(0)).max())
    return err


def test_face_frames_are_the_3d_projections():
    assert frame_errors(panel()) < 1e-13


def test_face_frame_check_fails_on_planted_mistake():
    with O.plant("gno_flux_cos"):
        assert frame_errors(panel()) > 1e-3


def rest_tendency(riemann):
    c = panel(riemann=riemann)
    n3, n2, n1 = O.shape(c)
    w = np.zeros((5, n3, n2, n1))
    w[0], w[4] = 1.3, 0.9
    w, u = O.initialize(c, w)
    du, _, _ = O.hydro_forward(c, 0.01, u)
    I = O.interior(c)
    dx = (c.x1max - c.x1min) / c.nx1
    return max(np.abs(du[v][I[1:]]).max() for v in range(5)) / (0.01 * 0.9 / dx)


@pytest.mark.parametrize("riemann", ["lmars", "hllc"])
def test_uniform_pressure_at_rest_has_no_tendency(riemann):
    assert rest_tendency(riemann) < 1e-13


@pytest.mark.parametrize("mistake", ["gno_flux_cos", "gno_xov_sin"])
def test_rest_check_fails_on_planted_mistake(mistake):
    with O.plant(mistake):
        assert rest_tendency("lmars") > 1e-4


def test_roe_takes_no_face_frame():
    """The reference's Roe solver does not rotate to the face frame
    (src/riemann/roe.cpp:14-59), so a uniform pressure at rest on a panel is
    not in balance with it: its x2 / x3 momentum fluxes carry p where the
    metric sources expect p sin(theta_face)."""
    assert rest_tendency("roe") > 1e-3


@pytest.mark.parametrize("scheme", [0, 1])
def test_rk3_conserves_mass_with_reflecting_walls(scheme):
    """A perturbation at rest in the middle of a closed panel under gravity:
    one step reaches nine cells (three stages of a three-cell stencil), so
    the flow does not reach the side walls, and the radial walls reflect
    exactly.  (At a side wall the reference's reflection flips the covariant
    momentum, which does not zero the face's normal velocity where the
    coordinate lines are not orthogonal; a flow reaching the side walls
    leaks mass through them.)"""
    c = panel(nx1=6, nx=24, riemann=("lmars", "hllc", "lmars"), bc=("reflecting",) * 6,
              grav1=-0.4, implicit_scheme=scheme)
    n3, n2, n1 = O.shape(c)
    w = np.zeros((5, n3, n2, n1))
    w[0], w[4] = 1.0, 1.0
    rng = np.random.default_rng(5)
    mid = slice(n2 // 2 - 2, n2 // 2 + 2)
    w[0][mid, mid] += rng.uniform(-0.1, 0.1, (4, 4, n1))
    w[4][mid, mid] += rng.uniform(-0.1, 0.1, (4, 4, n1))
    w, u = O.initialize(c, w)
    vol = O.geometry(c, "vol")
    I = O.interior(c)[1:]
    m0 = (vol * u[0])[I].sum()
    u, _ = O.rk3_step(c, 0.02, u)
    assert abs((vol * u[0])[I].sum() / m0 - 1) < 1e-14
