"""ctypes wrapper of the CPU oracle: TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this
module, and only as the checker.  The product library never loads it.

The C restatement (f4j_oracle.c) follows Flag4j's CPU matmul kernels; see the
citations there.  Arrays are numpy, row-major, contiguous; complex arrays are
numpy complex128 (interleaved re/im in memory, as the C code expects).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libf4j_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i = ctypes.c_int

KERNEL_NAMES = ["SMALL_STANDARD", "STRD", "REORD", "BLK", "BLK_REORD", "MT_STRD", "MT_REORD",
                "MT_BLK", "MT_BLK_REORD", "BLK_VEC", "MT_BLK_VEC", "MT_STRD_VEC"]
TRANSPOSE_NAMES = ["MULT_T", "MULT_T_BLOCKED", "MULT_T_CONCURRENT", "MULT_T_BLOCKED_CONCURRENT"]
OPS = {"N": 0, "T": 1, "C": 2}


class Thresholds(ctypes.Structure):
    _fields_ = [("small_threshold", ctypes.c_int), ("aspect_threshold", ctypes.c_double),
                ("sq_mt_standard", ctypes.c_int), ("sq_mt_reordered", ctypes.c_int),
                ("wide_mt_reordered", ctypes.c_int), ("sq_seq_vec", ctypes.c_int),
                ("min_dim_small", ctypes.c_int)]


# Defaults of M:linalg/ops/dispatch/configs/ReDeMatMultDispatchConfigs.java:52-86
DEFAULT_THRESHOLDS = dict(small_threshold=2450, aspect_threshold=2.0, sq_mt_standard=50,
                          sq_mt_reordered=2048, wide_mt_reordered=25, sq_seq_vec=256, min_dim_small=10)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name in ("f4o_real_standard", "f4o_real_reordered", "f4o_real_reordered_fma",
                     "f4o_real_mult_transpose"):
            getattr(L, name).argtypes = [_i, _i, _i, _dp, _dp, _dp]
        L.f4o_real_blocked_reordered.argtypes = [_i, _i, _i, _dp, _dp, _dp, _i]
        L.f4o_real_blocked_vector.argtypes = [_i, _i, _dp, _dp, _dp, _i]
        L.f4o_real_mt_reordered.argtypes = [_i, _i, _i, _dp, _dp, _dp, _i, _i, _i]
        L.f4o_real_mt_blocked_reordered.argtypes = [_i, _i, _i, _dp, _dp, _dp, _i, _i, _i, _i]
        L.f4o_cplx_gemm.argtypes = [_i, _i, _i, _i, _i, _dp, _dp, _dp]
        L.f4o_mixed_gemm.argtypes = [_i, _i, _i, _i, _dp, _dp, _dp]
        L.f4o_cplx_mt_reordered.argtypes = [_i, _i, _i, _dp, _dp, _dp, _i, _i, _i]
        L.f4o_real_mt_standard.argtypes = [_i, _i, _i, _dp, _dp, _dp, _i, _i, _i]
        L.f4o_cplx_mt_standard.argtypes = [_i, _i, _i, _dp, _dp, _dp, _i, _i, _i]
        L.f4o_select_real.argtypes = [_i, _i, _i, ctypes.POINTER(Thresholds)]
        L.f4o_select_cplx.argtypes = [_i, _i, _i, ctypes.POINTER(Thresholds)]
        L.f4o_select_transpose.argtypes = [_i]
        L.f4o_real_entries.argtypes = [_i, _i, _i, _i, _i, _dp, _dp, _i, _i64p, _i64p, _dp]
        L.f4o_cplx_entries.argtypes = [_i, _i, _i, _i, _i, _dp, _dp, _i, _i64p, _i64p, _dp]
        L.f4o_qr.argtypes = [_i, _i, _dp, _dp, _dp, _dp]
        L.f4o_lstsq.argtypes = [_i, _i, _i, _dp, _dp, _dp]
        L.f4o_real_dot2.argtypes = [_i, _i, _i, _dp, _dp, _i, _i, _i64p, _i64p, _dp, _dp, _dp]
        L.f4o_cplx_dot2.argtypes = [_i, _i, _i, _i, _i, _dp, _dp, _i, _i64p, _i64p,
                                    _dp, _dp, _dp, _dp, _dp]
        L.f4o_lcg_fill_uniform.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp]

        class Lcg(ctypes.Structure):
            _fields_ = [("seed", ctypes.c_uint64), ("hng", ctypes.c_int), ("ng", ctypes.c_double)]
        L.Lcg = Lcg
        L.f4o_lcg_init.argtypes = [ctypes.POINTER(Lcg), ctypes.c_int64]
        L.f4o_lcg_next_int.argtypes = [ctypes.POINTER(Lcg)]
        L.f4o_lcg_next_int.restype = ctypes.c_int32
        L.f4o_lcg_next_double.argtypes = [ctypes.POINTER(Lcg)]
        L.f4o_lcg_next_double.restype = ctypes.c_double
        L.f4o_lcg_next_gaussian.argtypes = [ctypes.POINTER(Lcg)]
        L.f4o_lcg_next_gaussian.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


# ---- real ----
def real_standard(A, B):
    A, B = _f64(A), _f64(B)
    M, K = A.shape
    N = B.shape[1]
    C = np.empty((M, N))
    lib().f4o_real_standard(M, N, K, _p(A), _p(B), _p(C))
    return C


def real_reordered(A, B):
    A, B = _f64(A), _f64(B)
    M, K = A.shape
    N = B.shape[1]
    C = np.empty((M, N))
    lib().f4o_real_reordered(M, N, K, _p(A), _p(B), _p(C))
    return C


def real_reordered_fma(A, B):
    A, B = _f64(A), _f64(B)
    M, K = A.shape
    N = B.shape[1]
    C = np.empty((M, N))
    lib().f4o_real_reordered_fma(M, N, K, _p(A), _p(B), _p(C))
    return C


def real_blocked_reordered(A, B, bs=64):
    A, B = _f64(A), _f64(B)
    M, K = A.shape
    N = B.shape[1]
    C = np.empty((M, N))
    lib().f4o_real_blocked_reordered(M, N, K, _p(A), _p(B), _p(C), bs)
    return C


def real_blocked_vector(A, x, bs=64):
    A, x = _f64(A), _f64(x).reshape(-1)
    M, K = A.shape
    y = np.empty(M)
    lib().f4o_real_blocked_vector(M, K, _p(A), _p(x), _p(y), bs)
    return y


def real_mult_transpose(A, Bt):
    """C = A @ Bt.T with Bt stored N x K."""
    A, Bt = _f64(A), _f64(Bt)
    M, K = A.shape
    N = Bt.shape[0]
    C = np.empty((M, N))
    lib().f4o_real_mult_transpose(M, N, K, _p(A), _p(Bt), _p(C))
    return C


def real_mt_reordered(A, B, C=None, rows=None, threads=1):
    A, B = _f64(A), _f64(B)
    M, K = A.shape
    N = B.shape[1]
    if C is None:
        C = np.empty((M, N))
    r0, r1 = rows if rows is not None else (0, M)
    lib().f4o_real_mt_reordered(M, N, K, _p(A), _p(B), _p(C), r0, r1, threads)
    return C


def real_mt_blocked_reordered(A, B, C=None, rows=None, threads=1, bs=64):
    A, B = _f64(A), _f64(B)
    M, K = A.shape
    N = B.shape[1]
    if C is None:
        C = np.empty((M, N))
    r0, r1 = rows if rows is not None else (0, M)
    lib().f4o_real_mt_blocked_reordered(M, N, K, _p(A), _p(B), _p(C), r0, r1, threads, bs)
    return C


def real_mt_standard(A, B, C=None, rows=None, threads=1):
    A, B = _f64(A), _f64(B)
    M, K = A.shape
    N = B.shape[1]
    if C is None:
        C = np.empty((M, N))
    r0, r1 = rows if rows is not None else (0, M)
    lib().f4o_real_mt_standard(M, N, K, _p(A), _p(B), _p(C), r0, r1, threads)
    return C


# ---- complex ----
def _op_shape(X, op):
    r, c = X.shape
    return (r, c) if op == "N" else (c, r)


def mixed_gemm(A, B):
    """C = A @ B with one real and one complex factor, in RealFieldDenseMatMult's order."""
    mode = 1 if np.iscomplexobj(A) else 0
    A = _c128(A) if mode == 1 else _f64(A)
    B = _f64(B) if mode == 1 else _c128(B)
    M, K = A.shape
    N = B.shape[1]
    C = np.empty((M, N), dtype=np.complex128)
    lib().f4o_mixed_gemm(mode, M, N, K, _p(A.view(np.float64)), _p(B.view(np.float64)), _p(C.view(np.float64)))
    return C


def cplx_gemm(A, B, opA="N", opB="N"):
    A, B = _c128(A), _c128(B)
    M, K = _op_shape(A, opA)
    K2, N = _op_shape(B, opB)
    assert K == K2
    C = np.empty((M, N), dtype=np.complex128)
    lib().f4o_cplx_gemm(OPS[opA], OPS[opB], M, N, K, _p(A.view(np.float64)), _p(B.view(np.float64)),
                        _p(C.view(np.float64)))
    return C


def cplx_mt_reordered(A, B, C=None, rows=None, threads=1):
    A, B = _c128(A), _c128(B)
    M, K = A.shape
    N = B.shape[1]
    if C is None:
        C = np.empty((M, N), dtype=np.complex128)
    r0, r1 = rows if rows is not None else (0, M)
    lib().f4o_cplx_mt_reordered(M, N, K, _p(A.view(np.float64)), _p(B.view(np.float64)),
                                _p(C.view(np.float64)), r0, r1, threads)
    return C


def cplx_mt_standard(A, B, C=None, rows=None, threads=1):
    A, B = _c128(A), _c128(B)
    M, K = A.shape
    N = B.shape[1]
    if C is None:
        C = np.empty((M, N), dtype=np.complex128)
    r0, r1 = rows if rows is not None else (0, M)
    lib().f4o_cplx_mt_standard(M, N, K, _p(A.view(np.float64)), _p(B.view(np.float64)),
                               _p(C.view(np.float64)), r0, r1, threads)
    return C


# ---- dispatch ----
def _thr(overrides=None):
    d = dict(DEFAULT_THRESHOLDS)
    d.update(overrides or {})
    return Thresholds(**d)


def select_real(m, n, k, **overrides):
    t = _thr(overrides)
    return KERNEL_NAMES[lib().f4o_select_real(m, n, k, ctypes.byref(t))]


def select_cplx(m, n, k, **overrides):
    t = _thr(overrides)
    return KERNEL_NAMES[lib().f4o_select_cplx(m, n, k, ctypes.byref(t))]


def select_transpose(rows):
    return TRANSPOSE_NAMES[lib().f4o_select_transpose(rows)]


# ---- sampled entries in the reference's arithmetic ----
def real_entries(A, B, I, J, opA="N", opB="N"):
    """(op(A) op(B))_{I,J} as Flag4j's real kernels compute them (bitwise)."""
    A, B = _f64(A), _f64(B)
    M, K = A.shape if opA == "N" else A.shape[::-1]
    N = B.shape[1] if opB == "N" else B.shape[0]
    I = np.ascontiguousarray(I, dtype=np.int64)
    J = np.ascontiguousarray(J, dtype=np.int64)
    out = np.empty(len(I))
    lib().f4o_real_entries(OPS[opA], OPS[opB], M, N, K, _p(A), _p(B), len(I), I.ctypes.data_as(_i64p),
                           J.ctypes.data_as(_i64p), _p(out))
    return out


def cplx_entries(A, B, I, J, opA="N", opB="N"):
    """(op(A) op(B))_{I,J} as Flag4j's complex kernels compute them (bitwise)."""
    A, B = _c128(A), _c128(B)
    M, K = _op_shape(A, opA)
    _, N = _op_shape(B, opB)
    I = np.ascontiguousarray(I, dtype=np.int64)
    J = np.ascontiguousarray(J, dtype=np.int64)
    out = np.empty(len(I), dtype=np.complex128)
    lib().f4o_cplx_entries(OPS[opA], OPS[opB], M, N, K, _p(A.view(np.float64)), _p(B.view(np.float64)), len(I),
                           I.ctypes.data_as(_i64p), J.ctypes.data_as(_i64p), _p(out.view(np.float64)))
    return out


# ---- least squares ----
def qr(A):
    """(T, qf, Q): reduced Householder QR as RealQR computes it."""
    A = _f64(A)
    M, N = A.shape
    mn = min(M, N)
    T, qf, Q = np.empty((M, N)), np.empty(mn), np.empty((M, mn))
    lib().f4o_qr(M, N, _p(A), _p(T), _p(qf), _p(Q))
    return T, qf, Q


def lstsq(A, B):
    """(status, X): RealLstsqSolver.solve(A, B); status 0 ok, 1 singular R, 2 M < N."""
    A = _f64(A)
    B2 = _f64(B.reshape(B.shape[0], -1))
    M, N = A.shape
    R = B2.shape[1]
    X = np.empty((N, R))
    st = lib().f4o_lstsq(M, N, R, _p(A), _p(B2), _p(X))
    return st, X.reshape((N,) + B.shape[1:])


# ---- accuracy oracle ----
def real_dot2(A, B, I, J, transB=False):
    """Compensated (C)_{I,J}; returns (hi, lo, |A||B|)."""
    A, B = _f64(A), _f64(B)
    M, K = A.shape
    N = B.shape[0] if transB else B.shape[1]
    I = np.ascontiguousarray(I, dtype=np.int64)
    J = np.ascontiguousarray(J, dtype=np.int64)
    ns = len(I)
    hi, lo, ab = np.empty(ns), np.empty(ns), np.empty(ns)
    lib().f4o_real_dot2(M, N, K, _p(A), _p(B), int(transB), ns, I.ctypes.data_as(_i64p),
                        J.ctypes.data_as(_i64p), _p(hi), _p(lo), _p(ab))
    return hi, lo, ab


def cplx_dot2(A, B, I, J, opA="N", opB="N"):
    """Compensated (op(A) op(B))_{I,J}; returns (re_hi, re_lo, im_hi, im_lo, absab)."""
    A, B = _c128(A), _c128(B)
    M, K = _op_shape(A, opA)
    _, N = _op_shape(B, opB)
    I = np.ascontiguousarray(I, dtype=np.int64)
    J = np.ascontiguousarray(J, dtype=np.int64)
    ns = len(I)
    out = [np.empty(ns) for _ in range(5)]
    lib().f4o_cplx_dot2(OPS[opA], OPS[opB], M, N, K, _p(A.view(np.float64)), _p(B.view(np.float64)),
                        ns, I.ctypes.data_as(_i64p), J.ctypes.data_as(_i64p), *[_p(o) for o in out])
    return tuple(out)


# ---- java.util.Random ----
class JavaRandom:
    def __init__(self, seed):
        self._s = lib().Lcg()
        lib().f4o_lcg_init(ctypes.byref(self._s), seed)

    def next_int(self):
        return lib().f4o_lcg_next_int(ctypes.byref(self._s))

    def next_double(self):
        return lib().f4o_lcg_next_double(ctypes.byref(self._s))

    def next_gaussian(self):
        return lib().f4o_lcg_next_gaussian(ctypes.byref(self._s))


def lcg_uniform(seed, n):
    out = np.empty(n)
    lib().f4o_lcg_fill_uniform(seed, n, _p(out))
    return out
