"""snapy_linfel_amd: MI355X-native finite-volume RHS and vertically implicit
step of snapy (WENO5 / CP5 reconstruction, LMARS / HLLC / Roe fluxes,
ideal-gas EOS, Cartesian divergence, block-tridiagonal implicit solve,
SSP-RK3), with the reference's module API.

Native parts:
  libsnapx.so  hand-written HIP kernels (gfx950) behind the C ABI of
               include/snapx.h; bound here by ``capi`` (ctypes)
  _snapy*.so   C++ host modules (Hydro, Reconstruct, RiemannSolver,
               EquationOfState, ImplicitHydro, MeshBlock) over libtorch
"""
from . import capi  # noqa: F401

__all__ = ["capi", "host"]


def __getattr__(name):
    if name == "host":
        return _load_host()
    raise AttributeError(name)


def _load_host():
    """The C++ host modules (Hydro, Reconstruct, ...); raises if not built."""
    import importlib
    import torch  # noqa: F401  (libtorch must be loaded first)
    return importlib.import_module("snapy_linfel_amd._snapy")
