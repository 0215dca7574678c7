"""gfx950 GMRES for the 2-D Helmholtz Green's-function integral equation.

Layers:
  include/cugmres.h + cugmres_amd/csrc/*.hip  HIP kernels and the C ABI (libcugmres.so)
  cugmres_amd._lib                            ctypes binding of that ABI
  cugmres_amd.solver                          device-level API (torch tensors as memory)
  cugmres_amd.gmres                           the reference Python module's functions
  cugmres_amd.formats                         the reference's text file formats
  cugmres_amd.problems                        seeded permittivity maps and plane waves
"""
from . import problems  # noqa: F401

__all__ = ["problems", "solver", "gmres", "formats"]
__version__ = "0.1.0"
