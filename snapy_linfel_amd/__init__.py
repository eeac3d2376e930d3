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

__all__ = ["capi"]
