"""MI355X-native Chebyshev-filtered subspace iteration for LYNX.

``lynx_amd.core`` mirrors the pylynx classes on the eigensolver path
(Lattice, FDGrid, FDStencil, Domain, HaloExchange, NonlocalProjector,
Hamiltonian, EigenSolver) on top of the C ABI of ``liblynx_chefsi.so``.
"""
from . import _lib  # noqa: F401

__all__ = ["core", "build"]
