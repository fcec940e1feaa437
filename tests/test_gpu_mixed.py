"""GPU parity of the mixed real / complex products (f4j_dzgemm, f4j_zdgemm:
Matrix.mult(CMatrix / CVector), CMatrix.mult(Matrix / Vector)) against the oracle.

Bars:
  * STRICT, the small path and the matrix-vector forms (N = 1): bit for bit
    equal to the oracle's restatement of RealFieldDenseMatMult
    (RealFieldDenseMatMult.java:58-1090), and to the reference's golden vectors;
  * FP64 (MFMA): |C - C_ref| <= 2 K u (|A||B|) per component, |z| = |re| + |im|;
  * SPLIT_INT8: max |C - AB| / (|A||B|) <= 1e-15 on sampled entries against the
    compensated (Dot2) oracle.
"""
import ctypes

import numpy as np
import pytest

import flag4j_amd as f4j
from flag4j_amd import _native, device
from conftest import cplx_mat, real_mat
from oracle import oracle as o

pytestmark = pytest.mark.gpu
U = 2.0**-53


def wide(rng, *shape):
    return rng.standard_normal(shape) * 10.0 ** rng.integers(-6, 7, shape)


def wide_c(rng, *shape):
    return wide(rng, *shape) + 1j * wide(rng, *shape)


def p(a):
    return ctypes.c_void_p(a.ctypes.data)


def mixed(A, B, precision="auto", slices=0):
    M, K = A.shape
    N = B.shape[1]
    C = np.full((M, N), np.nan + 0j)
    f = _native.load().f4j_zdgemm if np.iscomplexobj(A) else _native.load().f4j_dzgemm
    _native.check(f(M, N, K, p(A), max(K, 1), p(B), max(N, 1), p(C), max(N, 1), _native.PREC[precision], slices))
    return C


def same_bits(X, Y):
    return np.array_equal(np.ascontiguousarray(X).view(np.float64), np.ascontiguousarray(Y).view(np.float64))


def absz(X):
    return np.abs(X.real) + np.abs(X.imag) if np.iscomplexobj(X) else np.abs(X)


@pytest.mark.parametrize("case", ["real_cmat_api_mult", "cplx_rmat_api_mult", "real_cvec_api_matvec",
                                  "cplx_vec_api_matvec"])
def test_golden_vectors_through_the_api(golden, case):
    """MatrixMultTests.java:66-87, CMatrixMultTests.java:20-41, MatrixVectorTests.java:62-83,
    CMatrixMatVecMultTests.java:22-45 through the mirrored API (CMatrix / CVector equality)."""
    g = golden[case]
    if case.startswith("cplx"):
        a = f4j.CMatrix(cplx_mat(g["a"]))
        b = f4j.Vector(real_mat(g["b"]).reshape(-1)) if "vec" in case else f4j.Matrix(real_mat(g["b"]))
    else:
        a = f4j.Matrix(real_mat(g["a"]))
        b = f4j.CVector(cplx_mat(g["b"])) if "vec" in case else f4j.CMatrix(cplx_mat(g["b"]))
    exp = cplx_mat(g["c"]).reshape(-1)
    c = a.mult(b)
    assert isinstance(c, f4j.CVector if "vec" in case else f4j.CMatrix)
    assert np.all((c.data.real == exp.real) & (c.data.imag == exp.imag))


@pytest.mark.parametrize("real_left", [True, False])
@pytest.mark.parametrize("M,N,K", [(37, 45, 130), (128, 64, 64), (5, 300, 7), (200, 3, 1000)])
def test_strict_bitwise(real_left, M, N, K):
    rng = np.random.default_rng(M + 7 * N + 13 * K)
    A = wide(rng, M, K) if real_left else wide_c(rng, M, K)
    B = wide_c(rng, K, N) if real_left else wide(rng, K, N)
    assert same_bits(mixed(A, B, "strict"), o.mixed_gemm(A, B))


@pytest.mark.parametrize("real_left", [True, False])
def test_small_path_bitwise(real_left):
    rng = np.random.default_rng(3)
    A = wide(rng, 20, 30) if real_left else wide_c(rng, 20, 30)
    B = wide_c(rng, 30, 40) if real_left else wide(rng, 30, 40)   # 600 + 1200 < 2450
    assert same_bits(mixed(A, B), o.mixed_gemm(A, B))


@pytest.mark.parametrize("real_left", [True, False])
@pytest.mark.parametrize("M,K", [(1000, 1000), (63, 4097), (64, 2), (33, 999), (1, 16)])
def test_matrix_vector_bitwise(real_left, M, K):
    """N = 1: Matrix.mult(CVector) / CMatrix.mult(Vector), one k-ascending chain per component,
    in FP64 and AUTO precision (odd K with a real matrix takes the strict kernel: same bits)."""
    rng = np.random.default_rng(M * 3 + K)
    A = wide(rng, M, K) if real_left else wide_c(rng, M, K)
    x = wide_c(rng, K, 1) if real_left else wide(rng, K, 1)
    ref = o.mixed_gemm(A, x)
    for prec in ("auto", "fp64", "strict"):
        assert same_bits(mixed(A, x, prec), ref), prec


def test_matrix_vector_through_api():
    rng = np.random.default_rng(8)
    A, x = wide(rng, 500, 800), wide_c(rng, 800)
    y = f4j.Matrix(A).mult(f4j.CVector(x))
    assert same_bits(y.data, o.mixed_gemm(A, x.reshape(-1, 1)).reshape(-1))
    Ac, xr = wide_c(rng, 300, 640), wide(rng, 640)
    y = f4j.CMatrix(Ac).mult(f4j.Vector(xr))
    assert same_bits(y.data, o.mixed_gemm(Ac, xr.reshape(-1, 1)).reshape(-1))


@pytest.mark.parametrize("real_left", [True, False])
@pytest.mark.parametrize("M,N,K", [(512, 512, 512), (300, 129, 1000), (1000, 64, 77)])
def test_fp64_within_bound(real_left, M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K)) if real_left else rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))
    B = (rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))) if real_left else rng.standard_normal((K, N))
    C = mixed(A, B, "fp64")
    ref = o.mixed_gemm(A, B)
    bound = 2 * K * U * (absz(A) @ absz(B))
    assert np.all(np.abs(C.real - ref.real) <= bound) and np.all(np.abs(C.imag - ref.imag) <= bound)


@pytest.mark.parametrize("real_left", [True, False])
def test_split_int8_accuracy(real_left):
    """The modular INT8 form on the real views reaches FP64 accuracy (<= 1e-15 of |A||B|)."""
    rng = np.random.default_rng(12)
    M, N, K = 384, 200, 700
    A = rng.random((M, K)) * 2 - 1 if real_left else (rng.random((M, K)) * 2 - 1) + 1j * (rng.random((M, K)) * 2 - 1)
    B = (rng.random((K, N)) * 2 - 1) + 1j * (rng.random((K, N)) * 2 - 1) if real_left else rng.random((K, N)) * 2 - 1
    C = mixed(A, B, "split_int8")
    I, J = rng.integers(0, M, 200), rng.integers(0, N, 200)
    rh, rl, ih, il, ab = o.cplx_dot2(A.astype(complex), B.astype(complex), I, J, "N", "N")
    err = np.maximum(np.abs((C[I, J].real - rh) - rl), np.abs((C[I, J].imag - ih) - il))
    assert np.max(err / ab) <= 1e-15


def test_device_leading_dimensions():
    """Device entries with padded leading dimensions (zd takes the stacked-planes path)."""
    rng = np.random.default_rng(21)
    M, N, K, lda, ldb, ldc = 150, 70, 333, 340, 75, 80
    L = _native.load()
    for real_left in (True, False):
        A = rng.standard_normal((M, lda)) if real_left else rng.standard_normal((M, lda)) + 1j * rng.standard_normal((M, lda))
        B = (rng.standard_normal((K, ldb)) + 1j * rng.standard_normal((K, ldb))) if real_left else rng.standard_normal((K, ldb))
        f = L.f4j_dzgemm_dev if real_left else L.f4j_zdgemm_dev
        with device.DeviceBuffer.from_array(A) as dA, device.DeviceBuffer.from_array(B) as dB, \
                device.DeviceBuffer(M * ldc * 16) as dC:
            dC.upload(np.full(M * ldc, 3.0 + 3.0j))
            for prec in ("strict", "fp64"):
                _native.check(f(M, N, K, dA.ptr, lda, dB.ptr, ldb, dC.ptr, ldc, _native.PREC[prec], 0, None))
                device.sync()
                C = dC.download(np.complex128, M * ldc).reshape(M, ldc)
                ref = o.mixed_gemm(np.ascontiguousarray(A[:, :K]), np.ascontiguousarray(B[:, :N]))
                if prec == "strict":
                    assert same_bits(C[:, :N], ref)
                else:
                    bound = 2 * K * U * (absz(A[:, :K]) @ absz(B[:, :N]))
                    assert np.all(np.abs(C[:, :N].real - ref.real) <= bound)
                    assert np.all(np.abs(C[:, :N].imag - ref.imag) <= bound)
                assert np.all(C[:, N:] == 3.0 + 3.0j)


def test_matrix_vector_alignment():
    """N = 1 on device operands at 8-byte (not 16-byte) aligned addresses: a real
    matrix with odd K and odd row length runs the chain kernel (it reads real
    entries as single doubles), and a complex vector or matrix that is not
    16-byte aligned takes a path that reads it safely.  Every result equals the
    reference's arithmetic bit for bit."""
    rng = np.random.default_rng(33)
    L = _native.load()
    M, K = 257, 333
    for real_left in (True, False):
        A = wide(rng, M, K) if real_left else wide_c(rng, M, K)
        x = wide_c(rng, K, 1) if real_left else wide(rng, K, 1)
        ref = o.mixed_gemm(A, x)
        f = L.f4j_dzgemm_dev if real_left else L.f4j_zdgemm_dev
        for a_off, x_off in ((0, 0), (8, 0), (0, 8), (8, 8)):
            pa = np.concatenate([np.zeros(a_off // 8), A.reshape(-1).view(np.float64)])
            px = np.concatenate([np.zeros(x_off // 8), x.reshape(-1).view(np.float64)])
            with device.DeviceBuffer.from_array(pa) as dA, device.DeviceBuffer.from_array(px) as dx, \
                    device.DeviceBuffer(M * 16) as dy:
                dy.upload(np.full(M, np.nan + 0j))
                _native.check(f(M, 1, K, dA.ptr + a_off, K, dx.ptr + x_off, 1, dy.ptr, 1, _native.PREC["fp64"], 0,
                                None))
                device.sync()
                y = dy.download(np.complex128, M).reshape(M, 1)
                assert same_bits(y, ref), (real_left, a_off, x_off)


def test_empty_and_zero_inner():
    C = mixed(np.zeros((4, 0)), np.zeros((0, 3), dtype=complex))
    assert np.all(C == 0) and C.shape == (4, 3)
    C = mixed(np.zeros((0, 5), dtype=complex), np.zeros((5, 3)))
    assert C.shape == (0, 3)
