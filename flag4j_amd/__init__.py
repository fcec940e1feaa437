"""flag4j_amd: MI355X-native backend for Flag4j's dense matrix multiply.

The GPU work is done by libflag4j_amd.so (HIP kernels for gfx950 behind the C
ABI in include/flag4j_amd.h).  This package mirrors the reference's Java API
for the multiply path (Matrix, CMatrix, Vector, CVector, Shape), its least-squares and pseudo-inverse callers, and
exposes the C ABI to Python.  Importing it loads the native library and fails if it is missing.
"""
from . import _native
from .arrays import (CMatrix, CVector, IllegalArgumentException, Invert, LinearAlgebraException, Matrix, RealLstsqSolver,
                     Shape, SingularMatrixException, Vector)
from .device import DeviceBuffer, dgemm_dev, sync, zgemm_dev

_native.load()

__all__ = ["Matrix", "CMatrix", "Vector", "CVector", "Shape", "LinearAlgebraException", "IllegalArgumentException",
           "RealLstsqSolver", "Invert", "SingularMatrixException",
           "DeviceBuffer", "dgemm_dev", "zgemm_dev", "sync"]
