"""The differentiable PyTorch restatement of the step (oracle/torch_oracle.py)
against the C oracle: the same step to 1e-12 per conserved variable (5x5
column solve: 1e-11), and torch autograd through it against central
differences of the C oracle.  CPU only."""
import numpy as np
import pytest
import torch

import oracle as O
import torch_oracle as T


def _straka_oracle(scheme, nx2=24, nx1=16, wind=10.0):
    oc = O.straka_config(nx2=nx2, nx1=nx1, implicit_scheme=scheme)
    oc.x2max = 25.6e3 * nx2 / 512 * 2
    w = O.straka_init(oc)
    n3, n2, n1 = O.shape(oc)
    z = (np.arange(n1) - 3 + 0.5) / nx1
    x = (np.arange(n2) - 3 + 0.5) / nx2
    X, Z = np.meshgrid(x, z, indexing="ij")
    w[1, 0] = wind * np.sin(np.pi * Z) * np.sin(2 * np.pi * X)
    w[2, 0] = wind * np.cos(np.pi * Z) * np.cos(2 * np.pi * X)
    w, u = O.initialize(oc, w)
    return oc, u, 0.2


def _box3d_oracle(riemann=("lmars", "roe", "roe"), n=(8, 6, 5)):
    oc = O.make_config(nx1=n[0], nx2=n[1], nx3=n[2], nghost=3, x1min=0.0, x1max=800.0,
                       x2min=0.0, x2max=600.0, x3min=0.0, x3max=500.0, gammad=1.4,
                       weight=8.31446261815324 / 287.0, recon_vertical=("weno5", False, False),
                       riemann=riemann, grav1=-9.81, implicit_scheme=1,
                       bc=("reflecting", "reflecting", "periodic", "periodic", "periodic",
                           "periodic"))
    n3, n2, n1 = O.shape(oc)
    z = (np.arange(n1) - 3 + 0.5) * 100.0
    Tk = 250.0 - 9.81 * z / 1004.5
    p = 1e5 * (Tk / 250.0) ** (1004.5 / 287.0)
    rng = np.random.default_rng(1234)
    w = np.zeros((5, n3, n2, n1))
    w[4] = p
    w[0] = p / (287.0 * (Tk + rng.uniform(-0.5, 0.5, size=(n3, n2, n1))))
    w[1:4] = rng.uniform(-1, 1, size=(3, n3, n2, n1))
    w, u = O.initialize(oc, w)
    return oc, u, 0.2


def _close(a, b):
    scale = np.abs(b).reshape(5, -1).max(axis=1)
    scale[1:4] = np.maximum(scale[1:4], np.abs(b[0]).max() * 300.0)
    return (np.abs(a - b).reshape(5, -1).max(axis=1) / scale).max()


CASES = ["straka0", "straka1", "straka9", "box_lmars_roe", "box_hllc", "sod"]


def _case(name):
    if name.startswith("straka"):
        return _straka_oracle(int(name[6]))
    if name == "box_hllc":
        return _box3d_oracle(riemann="hllc")
    if name == "box_lmars_roe":
        return _box3d_oracle()
    oc = O.sod_config(48)
    w = O.sod_init(oc)
    x = (np.arange(O.shape(oc)[2]) - 3 + 0.5) / 48
    w[1] = 0.1 + 0.3 * np.sin(2 * np.pi * x)
    w, u = O.initialize(oc, w)
    return oc, u, O.max_time_step(oc, w, 0.5)


@pytest.mark.parametrize("case", CASES)
def test_torch_step_matches_c_oracle(case):
    oc, u, dt = _case(case)
    ref, _ = O.rk3_step(oc, dt, u)
    got = T.rk3_step(oc, dt, torch.from_numpy(u)).numpy()
    tol = 1e-11 if case == "straka9" else 1e-12
    assert _close(got, ref) < tol, (case, _close(got, ref))


def test_torch_autograd_matches_central_differences():
    """Autograd through the torch step equals central differences of the C
    oracle along random directions (pins the torch restatement as a
    differentiable oracle; Richardson-extrapolated differences)."""
    oc, u, dt = _case("box_lmars_roe")
    rng = np.random.default_rng(3)
    scale = np.abs(u).reshape(5, -1).max(axis=1).reshape(5, 1, 1, 1)
    v = rng.standard_normal(u.shape) * scale
    v = O.apply_bc(oc, v, O.CONSERVED)
    lam = rng.standard_normal(u.shape) / scale
    x = torch.from_numpy(u.copy()).requires_grad_(True)
    y = T.rk3_step(oc, dt, x)
    (g,) = torch.autograd.grad((y * torch.from_numpy(lam)).sum(), x)
    ad = float((g.numpy() * v).sum())

    def central(eps):
        a, _ = O.rk3_step(oc, dt, u + eps * v)
        b, _ = O.rk3_step(oc, dt, u - eps * v)
        return float(((a - b) * lam).sum() / (2 * eps))

    fd = (4 * central(1e-6) - central(2e-6)) / 3
    assert abs(ad - fd) / abs(fd) < 1e-6, (ad, fd)


def test_check_redo_known_answers():
    """Hand-set states with verdicts fixed by the reference's rules: interior
    only, NaN hides a negative density (min is NaN), the pressure test needs
    the energy screen, and E <= 0 with rho > 0 always means p <= 0."""
    oc, u, _ = _straka_oracle(0, nx2=8, nx1=8)
    u = torch.from_numpy(u)
    assert T.check_redo(oc, u) == 0
    x = u.clone(); x[0, 0, 5, 1] = -1.0          # ghost cell along x1
    assert T.check_redo(oc, x) == 0
    x = u.clone(); x[0, 0, 5, 6] = 0.0
    assert T.check_redo(oc, x) == 1
    x[0, 0, 7, 7] = float("nan")
    assert T.check_redo(oc, x) == 4
    x = u.clone(); x[4, 0, 4, 4] = -1.0
    assert T.check_redo(oc, x) == 2
    x = u.clone(); x[1, 0, 4, 4] = 1e6             # ke > E, E > 0: screen passes
    assert T.check_redo(oc, x) == 0
    x[4, 0, 6, 6] = 0.0                             # screen trips; p < 0 at (4, 4)
    assert T.check_redo(oc, x) == 2
    x = u.clone(); x[2, 0, 3, 3] = float("inf")
    assert T.check_redo(oc, x) == 4
