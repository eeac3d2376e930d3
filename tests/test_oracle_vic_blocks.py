"""Pins the oracle's implicit block assembly and Roe's supersonic / LLF branches
independently of the formulas' text.

Every check derives its expectation from the mathematics, not from the
reference's expressions:

* the flux Jacobian is the complex-step derivative of the textbook x1 Euler
  flux F(U);
* the Roe eigensystem satisfies R R^-1 = I, R diag(lambda) R^-1 equals the
  Euler Jacobian at the textbook Roe state (Roe 1981), Roe's property
  A(U_L, U_R)(U_R - U_L) = F(U_R) - F(U_L) holds, and |A| equals A sign(A)
  by Sylvester's formula over the three distinct eigenvalues u - c, u, u + c
  (Higham, Functions of Matrices, ch. 1);
* the blocks a, b, c of a column are I/dt plus the complex-step Jacobian of
  the first-order Roe-split flux divergence with |A| frozen, minus the
  gravity source Jacobian, with the reflecting ghosts as mirror images of the
  end cells (src/implicit/vic_assemble_partial_impl.h:28-111,
  src/implicit/vic_assemble_full_impl.h:15-106);
* a supersonic Roe face returns the Euler flux of the upwind state (by Roe's
  property the Roe formula gives it too in exact arithmetic; the switch makes
  it exact to rounding) and a
  face whose linearised intermediate density is negative returns the Rusanov
  flux (src/riemann/roe_impl.h:178-211; Einfeldt et al. 1991, J. Comput.
  Phys. 92, 273, for the negative intermediate density of strong
  rarefactions).

Each check is also run with one plausible mistake planted in the oracle and
must then fail.
"""
import numpy as np
import pytest

import oracle as O

GAMMA = 1.4
GM1 = GAMMA - 1.0
SUB3 = [0, 1, 4]


def cons(w):
    rho, v, p = w[0], w[1:4], w[4]
    return np.concatenate([[rho], rho * v, [p / GM1 + 0.5 * rho * (v @ v)]])


def euler_flux(U, n=1):
    """Euler flux along axis n (1, 2, 3 = index of the normal momentum)."""
    rho = U[0]
    p = GM1 * (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / rho)
    un = U[n] / rho
    F = U * un
    F[n] = F[n] + p
    F[4] = (U[4] + p) * un
    return F


def cs_jacobian(f, x, h=1e-30):
    """Complex-step Jacobian: exact to rounding for real-analytic f."""
    x = np.asarray(x, dtype=complex)
    cols = []
    for k in range(x.size):
        e = np.zeros(x.size, dtype=complex)
        e[k] = 1j * h
        cols.append(np.imag(f(x + e)) / h)
    return np.stack(cols, axis=1)


def roe_cons_state(wl, wr):
    """Textbook Roe average (sqrt(rho)-weighted u and H) as a conserved state,
    with its normal velocity and sound speed."""
    sl, sr = np.sqrt(wl[0]), np.sqrt(wr[0])
    Ul, Ur = cons(wl), cons(wr)
    Hl, Hr = (Ul[4] + wl[4]) / wl[0], (Ur[4] + wr[4]) / wr[0]
    v = (sl * wl[1:4] + sr * wr[1:4]) / (sl + sr)
    H = (sl * Hl + sr * Hr) / (sl + sr)
    rho = sl * sr
    p = rho * GM1 / GAMMA * (H - 0.5 * (v @ v))
    c = np.sqrt(GM1 * (H - 0.5 * (v @ v)))
    return cons(np.concatenate([[rho], v, [p]])), v[0], c


def matrix_abs(A, u, c):
    """|A| by Sylvester's formula over the distinct eigenvalues u - c, u, u + c
    of the (diagonalisable) Euler Jacobian A: |A| = sum_k |l_k| prod_{j != k}
    (A - l_j I) / (l_k - l_j)."""
    lam = [u - c, u, u + c]
    out = np.zeros_like(A)
    I = np.eye(A.shape[0])
    for k, lk in enumerate(lam):
        P = I.copy()
        for j, lj in enumerate(lam):
            if j != k:
                P = P @ (A - lj * I) / (lk - lj)
        out += abs(lk) * P
    return out


def rand_prim(rng, vmin=0.05, vmax=0.5):
    v = rng.uniform(vmin, vmax, 3) * rng.choice([-1.0, 1.0], 3)
    return np.array([rng.uniform(0.5, 2.0), *v, rng.uniform(0.5, 2.0)])


# ---------------------------------------------------------------- Jacobians


def flux_jacobian_error(seed=0, n=50):
    rng = np.random.default_rng(seed)
    err = 0.0
    for _ in range(n):
        w = rand_prim(rng)
        J = O.flux_jacobian5(w, GAMMA)
        ref = cs_jacobian(euler_flux, cons(w))
        err = max(err, np.abs(J - ref).max() / np.abs(ref).max())
    return err


def test_flux_jacobian_is_the_euler_jacobian():
    assert flux_jacobian_error() < 1e-13


def test_flux_jacobian_check_fails_on_planted_mistake():
    with O.plant("fluxjac"):
        assert flux_jacobian_error() > 1e-4


def eigensystem_errors(seed=1, n=50):
    rng = np.random.default_rng(seed)
    e = dict(ident=0.0, jac=0.0, roe=0.0, absA=0.0)
    for _ in range(n):
        wl = rand_prim(rng)
        wr = rand_prim(rng)
        R, Ri, lam, absA = O.roe_eigensystem(wl, wr, GAMMA)
        e["ident"] = max(e["ident"], np.abs(R @ Ri - np.eye(5)).max())
        A = R @ np.diag(lam) @ Ri
        Uhat, uhat, chat = roe_cons_state(wl, wr)
        Ahat = cs_jacobian(euler_flux, Uhat)
        scale = np.abs(Ahat).max()
        e["jac"] = max(e["jac"], np.abs(A - Ahat).max() / scale)
        dU = cons(wr) - cons(wl)
        dF = euler_flux(cons(wr)) - euler_flux(cons(wl))
        e["roe"] = max(e["roe"], np.abs(A @ dU - dF).max() / np.abs(dF).max())
        e["absA"] = max(e["absA"], np.abs(absA - matrix_abs(Ahat, uhat, chat)).max() / scale)
    return e


def test_roe_eigensystem():
    e = eigensystem_errors()
    assert e["ident"] < 1e-13, e
    assert e["jac"] < 1e-13, e
    assert e["roe"] < 1e-12, e
    assert e["absA"] < 1e-12, e


@pytest.mark.parametrize("mistake", ["rinv_sign", "roe_velocity"])
def test_roe_eigensystem_check_fails_on_planted_mistake(mistake):
    with O.plant(mistake):
        e = eigensystem_errors()
    assert max(e.values()) > 1e-3, e


# ------------------------------------------------------------- column blocks


def column_case(N, nl, seed, spherical):
    rng = np.random.default_rng(seed)
    w = np.stack([rand_prim(rng) for _ in range(nl + 2)], axis=1)
    # reflecting ends: the ghost is the mirror image of the end cell
    w[:, 0] = w[:, 1]
    w[1, 0] = -w[1, 1]
    w[:, -1] = w[:, -2]
    w[1, -1] = -w[1, -2]
    if spherical:
        rf = np.linspace(1.0, 2.0, nl + 1)
        area = rf ** 2
        vol = (rf[1:] ** 3 - rf[:-1] ** 3) / 3.0
    else:
        area = np.full(nl + 1, 0.7)
        vol = 0.7 * np.full(nl, 0.25)
    return w, area, vol


def linearised_blocks(w, area, vol, dt, grav):
    """I/dt + d/dU [div(first-order Roe-split flux, |A| frozen) - gravity]."""
    nl = w.shape[1] - 2
    Bnd = np.diag([1.0, -1.0, 1.0, 1.0, 1.0])
    absA = []
    for f in range(nl + 1):
        Uhat, uhat, chat = roe_cons_state(w[:, f], w[:, f + 1])
        absA.append(matrix_abs(cs_jacobian(euler_flux, Uhat), uhat, chat))
    Phi = np.zeros((5, 5))
    Phi[1, 0] = grav
    Phi[4, 1] = grav

    def residual(x):
        U = x.reshape(nl, 5)
        Ue = [Bnd @ U[0]] + [U[i] for i in range(nl)] + [Bnd @ U[-1]]
        F = [euler_flux(u) for u in Ue]
        Fh = [0.5 * (F[f] + F[f + 1]) - 0.5 * absA[f] @ (Ue[f + 1] - Ue[f])
              for f in range(nl + 1)]
        out = [(area[i + 1] * Fh[i + 1] - area[i] * Fh[i]) / vol[i] - Phi @ U[i]
               for i in range(nl)]
        return np.concatenate(out)

    x0 = np.concatenate([cons(w[:, i + 1]) for i in range(nl)])
    M = cs_jacobian(residual, x0) + np.eye(5 * nl) / dt
    blk = lambda i, j: M[5 * i:5 * i + 5, 5 * j:5 * j + 5]
    a = [blk(i, i) for i in range(nl)]
    b = [blk(i, i - 1) if i > 0 else np.zeros((5, 5)) for i in range(nl)]
    c = [blk(i, i + 1) if i < nl - 1 else np.zeros((5, 5)) for i in range(nl)]
    return np.array(a), np.array(b), np.array(c)


def block_error(N, nl, seed, spherical, dt=0.3, grav=-0.7):
    w, area, vol = column_case(N, nl, seed, spherical)
    a, b, c = O.vic_assemble(N, w, area, vol, dt, grav, GAMMA)
    ra, rb, rc = linearised_blocks(w, area, vol, dt, grav)
    if N == 3:
        ra, rb, rc = (x[:, SUB3][:, :, SUB3] for x in (ra, rb, rc))
    scale = max(np.abs(ra).max(), np.abs(rb).max(), np.abs(rc).max())
    # b at the bottom and c at the top hold the ghost couplings, which the
    # fold moves into a; the oracle keeps them, the linearisation has none
    err = max(np.abs(a - ra).max(), np.abs(b[1:] - rb[1:]).max(),
              np.abs(c[:-1] - rc[:-1]).max())
    return err / scale


CASES = [(N, nl, sph) for N in (3, 5) for nl in (4, 5, 6) for sph in (False, True)]


@pytest.mark.parametrize("N,nl,spherical", CASES)
def test_blocks_are_the_linearised_divergence(N, nl, spherical):
    assert block_error(N, nl, seed=10 * nl + N, spherical=spherical) < 1e-13


@pytest.mark.parametrize("mistake", ["c_jac_i", "no_bottom_fold", "phi_place", "b_abs_plus",
                                     "rinv_sign", "fluxjac"])
@pytest.mark.parametrize("N", [3, 5])
def test_block_check_fails_on_planted_mistake(mistake, N):
    # the J_i term of a_i and the J_{i+1} term of c_i differ only where the
    # face areas differ, so the spherical column carries every mistake
    with O.plant(mistake):
        err = block_error(N, 5, seed=7, spherical=True)
    assert err > 1e-4, err


# ---------------------------------------------------- Roe branches (oracle)


def faces(kind, dim, n=64, seed=0):
    """[5, 1, 1, n] left/right primitive states of one regime.

    kind: 'right' (both sides supersonic to +n), 'left' (to -n), 'llf'
    (strong rarefaction: sides moving apart at 4 units, sound speeds <= 2.4),
    'sub' (subsonic, close states)."""
    rng = np.random.default_rng(seed + {"right": 0, "left": 1, "llf": 2, "sub": 3}[kind])
    ivx = 4 - dim
    wl = np.empty((5, 1, 1, n))
    wr = np.empty((5, 1, 1, n))
    for q in range(n):
        a = rand_prim(rng)
        b = a * (1.0 + 0.1 * rng.uniform(-1.0, 1.0, 5))
        if kind == "right":
            a[ivx], b[ivx] = 4.0 + a[ivx], 4.0 + b[ivx]
        elif kind == "left":
            a[ivx], b[ivx] = -4.0 + a[ivx], -4.0 + b[ivx]
        elif kind == "llf":
            a[ivx], b[ivx] = -4.0 + a[ivx], 4.0 + b[ivx]
        wl[:, 0, 0, q], wr[:, 0, 0, q] = a, b
    return wl, wr


def expected_flux(kind, wl, wr, dim):
    ivx = 4 - dim
    out = np.empty_like(wl)
    for q in range(wl.shape[-1]):
        a, b = wl[:, 0, 0, q], wr[:, 0, 0, q]
        Fl, Fr = euler_flux(cons(a), ivx), euler_flux(cons(b), ivx)
        if kind == "right":
            out[:, 0, 0, q] = Fl
        elif kind == "left":
            out[:, 0, 0, q] = Fr
        else:
            smax = max(abs(a[ivx]) + np.sqrt(GAMMA * a[4] / a[0]),
                       abs(b[ivx]) + np.sqrt(GAMMA * b[4] / b[0]))
            out[:, 0, 0, q] = 0.5 * (Fl + Fr) - 0.5 * smax * (cons(b) - cons(a))
    return out


def roe_branch_error(kind, dim):
    wl, wr = faces(kind, dim)
    flx, _ = O.riemann("roe", wl, wr, dim, GAMMA)
    ref = expected_flux(kind, wl, wr, dim)
    scale = np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
    return (np.abs(flx - ref) / scale).max()


@pytest.mark.parametrize("kind", ["right", "left", "llf"])
@pytest.mark.parametrize("dim", [3, 2, 1])
def test_roe_branches(kind, dim):
    assert roe_branch_error(kind, dim) < 1e-14


@pytest.mark.parametrize("kind,mistake", [("right", "roe_upwind_swap"), ("left", "roe_upwind_swap"),
                                          ("llf", "roe_llf_speed")])
def test_roe_branch_check_fails_on_planted_mistake(kind, mistake):
    with O.plant(mistake):
        assert roe_branch_error(kind, 3) > 1e-3


def test_subsonic_faces_take_the_roe_formula():
    """The subsonic close states used as the fourth kind are neither upwind
    nor LLF: the flux differs from both one-sided Euler fluxes and from the
    Rusanov flux."""
    wl, wr = faces("sub", 3)
    flx, _ = O.riemann("roe", wl, wr, 3, GAMMA)
    for kind in ("right", "left", "llf"):
        ref = expected_flux(kind, wl, wr, 3)
        scale = np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
        assert (np.abs(flx - ref) / scale).max() > 1e-4
