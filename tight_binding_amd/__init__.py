"""MI355X-native kernel-polynomial (KPM) path of the tight_binding library.

Native libraries (built in-tree, see Makefile):
  libtbkpm.so   HIP engine for gfx950 -- include/tbkpm.h
  libtbhost.so  host model layer      -- include/tbhost.h

Python modules:
  lattice   UnitCell / PartialLattice and the sparse Hamiltonian builders
  kpm       device context, resident matrices and vector blocks, moments
  spectral  kernels (Jackson, Lorentz, ...), alpha_n and DoS reconstruction
"""
from . import _native  # noqa: F401
from .sparse import CSC  # noqa: F401

__all__ = ["CSC", "kpm", "lattice", "spectral"]
