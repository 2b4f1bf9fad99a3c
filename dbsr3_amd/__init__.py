"""dbsr3_amd -- MI355X (gfx950) engine for DBSR's two-electron B-spline
Slater integrals: R^k/S^k arrays, their contraction with orbital densities,
coefficient-list accumulation into Hamiltonian blocks, and the adjoint
(direct+exchange potentials) of E = sum c*R^k.

The compute path is libdbsr3_amd.so (hand-written HIP kernels behind a C-ABI,
include/dbsr3_amd.h).  This package is the host-side mirror of the reference
interface for Python callers and tests; it has no CPU fallback.
"""
from .capi import DBSError, check, lib  # noqa: F401
from .dbs import DBS  # noqa: F401
from .engine import Engine  # noqa: F401
from .cdata import CData, decode_int  # noqa: F401
from .matrix import MatBlocks  # noqa: F401

__all__ = ["DBS", "Engine", "CData", "decode_int", "MatBlocks", "DBSError", "lib", "check"]
