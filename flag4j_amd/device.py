"""Device-resident operands: library-owned device buffers and the *_dev GEMM calls.

Callers that keep operands on the GPU across many products (the batched
entry, benchmarks, a Java caller holding pinned/device buffers) use these
instead of the host-array calls, which stage through the library's pool.
Pointers may also come from another allocator (e.g. a torch tensor's
data_ptr()); `stream` is a hipStream_t handle as an int (0 = library stream).
"""
import ctypes

import numpy as np

from . import _native


class DeviceBuffer:
    """A device allocation made through f4j_dev_alloc, freed on close()."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _native.check(_native.load().f4j_dev_alloc(nbytes, ctypes.byref(p)))
        self.ptr, self.nbytes = p.value, nbytes

    @classmethod
    def from_array(cls, arr, stream=0):
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr, stream)
        return buf

    def upload(self, arr, stream=0):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _native.check(_native.load().f4j_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes, stream or None))
        _native.check(_native.load().f4j_stream_sync(stream or None))

    def download(self, dtype, count, stream=0):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _native.check(_native.load().f4j_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, stream or None))
        return out

    def close(self):
        if self.ptr:
            _native.check(_native.load().f4j_dev_free(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def dgemm_dev(opA, opB, M, N, K, A, lda, B, ldb, C, ldc, precision="fp64", slices=0, stream=0):
    """Kernel-level real GEMM on device pointers (ints); asynchronous on `stream`."""
    _native.check(_native.load().f4j_dgemm_dev(_native.OP[opA], _native.OP[opB], M, N, K, A, lda, B, ldb, C, ldc,
                                               _native.PREC[precision], slices, stream or None))


def zgemm_dev(opA, opB, M, N, K, A, lda, B, ldb, C, ldc, precision="fp64", slices=0, use_3m=False, stream=0):
    """Kernel-level complex128 GEMM on device pointers (ints); asynchronous on `stream`."""
    _native.check(_native.load().f4j_zgemm_dev(_native.OP[opA], _native.OP[opB], M, N, K, A, lda, B, ldb, C, ldc,
                                               _native.PREC[precision], slices, int(use_3m), stream or None))


def sync(stream=0):
    _native.check(_native.load().f4j_stream_sync(stream or None))
