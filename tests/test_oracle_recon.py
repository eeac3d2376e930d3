"""Pins the oracle's interpolants and Reconstruct composition.

Known answers come from the reference's own tests (tests/test_weno.cpp,
tests/test_plm.cpp); the rest are properties the mathematics fixes.
"""
import numpy as np
import pytest

import oracle as O


def test_weno5_known_answer():
    # tests/test_weno.cpp:175-184
    assert abs(O.interp_weno5(1.0, 2.0, 3.0, 4.0, 5.0) - 2.5000000000000004) < 1e-10


def test_weno3_known_answer():
    # tests/test_weno.cpp:145-152
    assert abs(O.interp_weno3(1.0, 2.0, 3.0) - 1.5) < 1e-10


def test_plm_known_answer():
    # tests/test_plm.cpp:20-27
    assert abs(O.interp_plm(1.0, 2.0, 3.0) - 1.5) < 1e-10


def test_weno3_smooth_limit():
    # tests/test_weno.cpp:154-173: WENO3 reduces to CP3 on smooth data
    eps = 1e-7
    phi = np.array([1 + eps, 1 + 2 * eps, 1 + 4 * eps])
    left = O.interp_window("weno3", 0, phi)
    right = O.interp_window("weno3", 1, phi)
    assert abs(left - O.interp_cp3(phi[0], phi[1], phi[2])) < 1e-6
    assert abs(right - O.interp_cp3(phi[2], phi[1], phi[0])) < 1e-6


@pytest.mark.parametrize("kind", ["weno5", "weno3", "cp5", "cp3"])
def test_table_form_matches_closed_form_and_mirror(kind):
    """The table-driven module kernels (interp_impl.h) equal the scalar
    closed forms (interp_simple.hpp) for left(); right() is left() on the
    mirrored window."""
    rng = np.random.default_rng(11)
    n = 5 if kind.endswith("5") else 3
    closed = {"weno5": O.interp_weno5, "weno3": O.interp_weno3,
              "cp5": O.interp_cp5, "cp3": O.interp_cp3}[kind]
    for _ in range(200):
        phi = rng.normal(size=n) * rng.choice([1e-3, 1.0, 1e5])
        left = O.interp_window(kind, 0, phi)
        right = O.interp_window(kind, 1, phi)
        scale = np.abs(phi).max()
        assert abs(left - closed(*phi)) <= 1e-14 * scale
        assert abs(right - closed(*phi[::-1])) <= 1e-14 * scale
        # the mirror identity holds for the table form itself
        assert abs(right - O.interp_window(kind, 0, phi[::-1])) <= 1e-14 * scale


def _cell_averages(k, n, dx=0.1):
    """Exact cell averages of x^k on cells [i dx, (i+1) dx]."""
    x = np.arange(n + 1) * dx
    return (x[1:] ** (k + 1) - x[:-1] ** (k + 1)) / ((k + 1) * dx), x


@pytest.mark.parametrize("kind,maxdeg", [("cp5", 4), ("weno5", 1), ("cp3", 2),
                                         ("weno3", 1)])
def test_polynomial_exactness(kind, maxdeg):
    """Face values reconstructed from exact cell averages of polynomials up to
    the scheme's degree are exact (CP5 deg 4, CP3 deg 2, WENO any sub-stencil
    combination exact for linear data)."""
    n = 5 if kind.endswith("5") else 3
    for k in range(maxdeg + 1):
        avg, x = _cell_averages(k, n)
        c = n // 2
        lo = O.interp_window(kind, 0, avg)
        hi = O.interp_window(kind, 1, avg)
        assert abs(lo - x[c] ** k) < 1e-12 * max(1.0, abs(x[c] ** k))
        assert abs(hi - x[c + 1] ** k) < 1e-12 * max(1.0, abs(x[c + 1] ** k))


def test_weno5_scale_option_invariance():
    """scale: true normalises by mean |phi|; the result is the unscaled value
    of the normalised window times the scale, and 0 for an all-zero window
    (src/recon/interp_impl.h:89-100)."""
    phi = np.array([1.0, 2.5, 2.0, 4.0, 3.0]) * 1e3
    s = np.abs(phi).mean()
    got = O.interp_window("weno5", 0, phi, scale=True)
    exp = O.interp_window("weno5", 0, phi / s) * s
    assert abs(got - exp) < 1e-12 * s
    assert O.interp_window("weno5", 0, np.zeros(5), scale=True) == 0.0


@pytest.mark.parametrize("dim", [3, 2, 1])
@pytest.mark.parametrize("shock", [True, False])
def test_reconstruct_dummy_fill_and_split(dim, shock):
    """ReconstructImpl::forward (src/recon/reconstruct.cpp:46-61,75-154):
    IRT holds lower-face values on faces [il, iu], ILT upper-face values of the
    cell below on faces [il, iu]; faces outside are copies of the edge faces;
    with shock: false, velocity and pressure use CP5."""
    rng = np.random.default_rng(5)
    shp = [5, 9, 10, 11]
    w = rng.uniform(0.5, 2.0, size=shp)
    wlr = O.reconstruct(w, dim, "weno5", shock=shock)
    ax = {3: 3, 2: 2, 1: 1}[dim]
    n = shp[ax]
    il, iu = 3, n - 3
    lt = np.moveaxis(wlr[0], ax, -1)
    rt = np.moveaxis(wlr[1], ax, -1)
    wm = np.moveaxis(w, ax, -1)
    for f in range(il, iu + 1):
        for v in range(5):
            kind = "weno5" if (shock or v == 0) else "cp5"
            # right state at face f: lower face of cell f
            exp_r = np.vectorize(lambda *p: O.interp_window(kind, 0, np.array(p)))(
                *[wm[v, ..., f - 2 + m] for m in range(5)])
            exp_l = np.vectorize(lambda *p: O.interp_window(kind, 1, np.array(p)))(
                *[wm[v, ..., f - 3 + m] for m in range(5)])
            assert np.array_equal(rt[v, ..., f], exp_r)
            assert np.array_equal(lt[v, ..., f], exp_l)
    for f in range(0, il):
        assert np.array_equal(rt[..., f], rt[..., il])
        assert np.array_equal(lt[..., f], lt[..., il])
    for f in range(iu + 1, n):
        assert np.array_equal(rt[..., f], rt[..., iu])
        assert np.array_equal(lt[..., f], lt[..., iu])


def test_reconstruct_constant_state_is_exact():
    w = np.ones((5, 1, 1, 12)) * np.array([1.2, 0.3, -0.1, 0.2, 7.0])[:, None, None, None]
    for kind in ["weno5", "weno3", "cp5", "cp3", "plm", "dc"]:
        wlr = O.reconstruct(w, 3, kind, shock=True)
        assert np.allclose(wlr[0], w, rtol=1e-15, atol=0)
        assert np.allclose(wlr[1], w, rtol=1e-15, atol=0)


def test_reconstruct_without_parent_clamps():
    # reconstruct.cpp:89-93: no parent Hydro -> interp1 everywhere, clamp_min(0)
    w = np.zeros((5, 1, 1, 12))
    w[1] = np.linspace(-1, 1, 12)
    wlr = O.reconstruct(w, 3, "weno5", has_parent=False)
    assert wlr.min() >= 0.0


def test_plm_module_form():
    # PLMInterpImpl::forward (src/recon/plm.cpp): monotone data -> harmonic slope
    w = np.zeros((5, 1, 1, 6))
    w[:] = np.array([1.0, 2.0, 3.0, 5.0, 4.0, 4.0])
    wlr = O.reconstruct(w, 3, "plm", shock=True)
    # cell 2 (2,3,5): harmonic slope
    assert abs(wlr[1, 0, 0, 0, 2] - O.interp_plm(2.0, 3.0, 5.0)) < 1e-14
    # extremum at cell 3 (3,5,4): slope zero
    assert wlr[1, 0, 0, 0, 3] == 5.0 and wlr[0, 0, 0, 0, 4] == 5.0
