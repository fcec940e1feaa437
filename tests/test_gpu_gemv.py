"""GPU parity of the matrix-vector kernels (f4j_dgemv / f4j_zgemv and the N = 1
route of f4j_dgemm / f4j_zgemm) against the oracle, bit for bit.

The reference's orders (RealDenseMatMult.java:426-569, DenseSemiringMatMult.java:435-574):
  * real, M K + K >= smallThreshold: 64-wide k blocks summed from +0.0, block sums
    added in order (oracle real_blocked_vector);
  * real small path and multTranspose with a one-row B: one k-ascending sum
    (oracle real_standard / real_mult_transpose);
  * complex: one k-ascending sum of Complex128 products (oracle cplx_gemm).
Inputs span many binades so that a different summation order changes the bits;
each test checks that its input separates the orders it must tell apart.
"""
import ctypes

import numpy as np
import pytest

import flag4j_amd as f4j
from flag4j_amd import _native, device
from conftest import bits, cplx_mat, real_mat
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def wide(rng, *shape):
    return rng.standard_normal(shape) * 10.0 ** rng.integers(-6, 7, shape)


def wide_c(rng, *shape):
    return wide(rng, *shape) + 1j * wide(rng, *shape)


def p(a):
    return ctypes.c_void_p(a.ctypes.data)


def dgemv(A, x, precision="auto"):
    M, K = A.shape
    y = np.full(M, np.nan)
    _native.check(_native.load().f4j_dgemv(M, K, p(A), max(K, 1), p(x), 1, p(y), 1, _native.PREC[precision]))
    return y


def zgemv(A, x, precision="auto"):
    M, K = A.shape
    y = np.full(M, np.nan + 0j)
    _native.check(_native.load().f4j_zgemv(M, K, p(A), max(K, 1), p(x), 1, p(y), 1, _native.PREC[precision]))
    return y


def test_golden_vectors_through_the_api(golden):
    """MatrixVectorTests.java:27-45 and CMatrixMatVecMultTests.java:62-84 through Matrix.mult(Vector)
    and CMatrix.mult(CVector)."""
    g = golden["real_api_matvec"]
    y = f4j.Matrix(real_mat(g["a"])).mult(f4j.Vector(real_mat(g["b"])))
    assert isinstance(y, f4j.Vector)
    assert np.array_equal(bits(y.data), bits(real_mat(g["c"])))
    g = golden["cplx_api_matvec"]
    y = f4j.CMatrix(cplx_mat(g["a"])).mult(f4j.CVector(cplx_mat(g["b"]).reshape(-1)))
    assert isinstance(y, f4j.CVector)
    assert np.array_equal(y.data.view(np.float64), cplx_mat(g["c"]).view(np.float64))
    g = golden["real_colvec"]                      # MatrixMultiplyTests.java:111-135 as an N = 1 product
    C = f4j.Matrix(real_mat(g["a"])).mult(f4j.Matrix(real_mat(g["b"])))
    assert np.array_equal(bits(C.data), bits(real_mat(g["c"])))


def test_real_blocked_exact_cases():
    """The blocked kernel on rows whose blocked-order results are fixed by hand in exact
    arithmetic (test_oracle.blocked_vector_cases): 40 copies of each row, sign
    alternating, so M K + K is above the small path and the blocked form runs."""
    from test_oracle import blocked_vector_cases
    for a, x, blocked, _ in blocked_vector_cases():
        sgn = np.where(np.arange(40) % 2, -1.0, 1.0)
        A = np.array([a] * 40) * sgn[:, None]
        y = dgemv(A, np.array(x))
        exp = sgn * blocked
        if blocked == 0.0:
            exp = np.zeros(40)          # +0.0 + (-0.0) = +0.0 in every row
        assert np.array_equal(bits(y), bits(exp)), (len(a), y[:2], blocked)


@pytest.mark.parametrize("M,K", [(1000, 1000), (257, 4097), (64, 512), (8, 513), (3, 1000), (100, 64),
                                 (100, 65), (9, 2001), (31, 1001), (2, 100000)])
def test_real_blocked_bitwise(M, K):
    rng = np.random.default_rng(M * 7919 + K)
    A, x = wide(rng, M, K), wide(rng, K)
    y = dgemv(A, x)
    ref = o.real_blocked_vector(A, x, 64)
    assert np.array_equal(bits(y), bits(ref)), (M, K)
    if K > 65:      # the input separates the blocked and plain orders (one element past a block adds alike)
        assert not np.array_equal(bits(ref), bits(o.real_standard(A, x.reshape(-1, 1)).reshape(-1)))


def test_real_small_path_is_one_chain():
    """M K + K < smallThreshold (2450): standardVector, one k-ascending sum per row."""
    rng = np.random.default_rng(11)
    A, x = wide(rng, 20, 100), wide(rng, 100)
    ref = o.real_standard(A, x.reshape(-1, 1)).reshape(-1)
    assert not np.array_equal(bits(ref), bits(o.real_blocked_vector(A, x, 64)))
    assert np.array_equal(bits(dgemv(A, x)), bits(ref))
    assert np.array_equal(bits(dgemv(A, x, "strict")), bits(ref))


@pytest.mark.parametrize("M,K", [(300, 700), (64, 64), (65, 1000), (1000, 3)])
def test_dgemm_column_vector_routes(M, K):
    """f4j_dgemm with N = 1: (N, N) large takes the blocked order in FP64 and STRICT
    precision (ReDeMatMultDispatcher.java:183-198); (N, T) is Matrix.multTranspose
    with a one-row B, one k-ascending sum (RealDenseMatMultTranspose.java:53-82)."""
    rng = np.random.default_rng(M + 3 * K)
    A, x = wide(rng, M, K), wide(rng, K)
    blocked = o.real_blocked_vector(A, x, 64)
    L = _native.load()
    for prec in ("auto", "strict", "fp64"):
        C = np.full(M, np.nan)
        _native.check(L.f4j_dgemm(0, 0, M, 1, K, p(A), K, p(x), 1, p(C), 1, _native.PREC[prec], 0))
        small = M * K + K < 2450
        ref = o.real_standard(A, x.reshape(-1, 1)).reshape(-1) if small else blocked
        assert np.array_equal(bits(C), bits(ref)), prec
        C = np.full(M, np.nan)
        _native.check(L.f4j_dgemm(0, 1, M, 1, K, p(A), K, p(x), K, p(C), 1, _native.PREC[prec], 0))
        assert np.array_equal(bits(C), bits(o.real_mult_transpose(A, x.reshape(1, -1)).reshape(-1))), prec


@pytest.mark.parametrize("M,K", [(500, 500), (33, 1000), (32, 16), (1, 7), (97, 4099), (1000, 1)])
def test_complex_bitwise(M, K):
    rng = np.random.default_rng(M * 31 + K)
    A, x = wide_c(rng, M, K), wide_c(rng, K)
    y = zgemv(A, x)
    ref = o.cplx_gemm(A, x.reshape(-1, 1)).reshape(-1)
    assert np.array_equal(y.view(np.float64), ref.view(np.float64)), (M, K)


@pytest.mark.parametrize("opB", ["N", "T", "C"])
def test_zgemm_column_vector_routes(opB):
    """f4j_zgemm with N = 1 (opB = C conjugates x) against the oracle's k-ascending sums."""
    rng = np.random.default_rng(5)
    M, K = 257, 900
    A, x = wide_c(rng, M, K), wide_c(rng, K)
    xs = x.reshape(-1, 1) if opB == "N" else x.reshape(1, -1)
    ref = o.cplx_gemm(A, xs, "N", opB).reshape(-1)
    for prec in ("auto", "strict", "fp64"):
        C = np.full(M, np.nan + 0j)
        _native.check(_native.load().f4j_zgemm(0, _native.OP[opB], M, 1, K, p(A), K, p(xs), xs.shape[1], p(C), 1,
                                               _native.PREC[prec], 0, 0))
        assert np.array_equal(C.view(np.float64), ref.view(np.float64)), prec


def test_device_strides_and_unaligned_rows():
    """Device entry with lda > K (odd: rows not 16-byte aligned), strided x and y:
    the register-staged form, still bitwise."""
    rng = np.random.default_rng(17)
    M, K, lda, incx, incy = 77, 1501, 1503, 3, 2
    A = wide(rng, M, lda)
    x = wide(rng, K * incx)
    L = _native.load()
    with device.DeviceBuffer.from_array(A) as dA, device.DeviceBuffer.from_array(x) as dx, \
            device.DeviceBuffer(M * incy * 8) as dy:
        dy.upload(np.full(M * incy, -7.0))
        _native.check(L.f4j_dgemv_dev(M, K, dA.ptr, lda, dx.ptr, incx, dy.ptr, incy, 0, None))
        device.sync()
        y = dy.download(np.float64, M * incy)
    ref = o.real_blocked_vector(np.ascontiguousarray(A[:, :K]), x[::incx][:K], 64)
    assert np.array_equal(bits(y[::incy]), bits(ref))
    assert np.all(y[1::incy] == -7.0)                 # untouched between strided results


def test_complex_device_strides():
    rng = np.random.default_rng(19)
    M, K, lda, incx, incy = 45, 333, 340, 2, 3
    A = wide_c(rng, M, lda)
    x = wide_c(rng, K * incx)
    L = _native.load()
    with device.DeviceBuffer.from_array(A) as dA, device.DeviceBuffer.from_array(x) as dx, \
            device.DeviceBuffer(M * incy * 16) as dy:
        dy.upload(np.full(M * incy, 5.0 + 5.0j))
        _native.check(L.f4j_zgemv_dev(M, K, dA.ptr, lda, dx.ptr, incx, dy.ptr, incy, 0, None))
        device.sync()
        y = dy.download(np.complex128, M * incy)
    ref = o.cplx_gemm(np.ascontiguousarray(A[:, :K]), x[::incx][:K].reshape(-1, 1)).reshape(-1)
    assert np.array_equal(np.ascontiguousarray(y[::incy]).view(np.float64), ref.view(np.float64))
    assert np.all(y[1::incy] == 5.0 + 5.0j)


def test_special_values_follow_the_reference():
    """Infinities, NaNs and signed zeros propagate exactly as in the oracle's order."""
    rng = np.random.default_rng(23)
    M, K = 200, 300
    A, x = wide(rng, M, K), wide(rng, K)
    A[3, 10] = np.inf
    A[4, 100] = -np.inf
    A[4, 200] = np.inf
    A[5, :] = -0.0
    x[250] = np.nan
    A[6, :] = 0.0
    y = dgemv(A, x)
    ref = o.real_blocked_vector(A, x, 64)
    assert np.array_equal(y, ref, equal_nan=True)
    fin = np.isfinite(ref)
    assert np.array_equal(bits(y[fin]), bits(ref[fin]))


def test_empty_and_zero_inner():
    assert dgemv(np.zeros((0, 5)), np.ones(5)).shape == (0,)
    y = dgemv(np.zeros((6, 0)), np.zeros(0))
    assert np.array_equal(bits(y), bits(np.zeros(6)))
    yc = zgemv(np.zeros((4, 0), dtype=complex), np.zeros(0, dtype=complex))
    assert np.array_equal(yc.view(np.float64), np.zeros(8))


def test_full_size_blocked_bitwise():
    """8192 x 8192 (the square configuration's operand), against the oracle."""
    rng = np.random.default_rng(29)
    M = K = 8192
    A = rng.random((M, K)) * 2 - 1
    x = rng.random(K) * 2 - 1
    assert np.array_equal(bits(dgemv(A, x)), bits(o.real_blocked_vector(A, x, 64)))


@pytest.mark.parametrize("M", [16385, 16383, 100])
def test_rows_per_wave_thresholds(M):
    """The one-chain-per-row kernels take 16 / 8 rows per wave (real A, by M) or 32 (complex
    A); each size class, with a ragged last wave, against the oracle bit for bit: Matrix.multTranspose with a one-row B (plain chain, K even: direct
    loads), CMatrix.mult(CVector), Matrix.mult(CVector) and CMatrix.mult(Vector)."""
    rng = np.random.default_rng(M)
    L = _native.load()
    K = 38
    A, x = wide(rng, M, K), wide(rng, K)
    C = np.full(M, np.nan)
    _native.check(L.f4j_dgemm(0, 1, M, 1, K, p(A), K, p(x), K, p(C), 1, _native.PREC["auto"], 0))
    assert np.array_equal(bits(C), bits(o.real_mult_transpose(A, x.reshape(1, -1)).reshape(-1)))
    xc = wide_c(rng, K)
    yc = np.full(M, np.nan + 0j)
    _native.check(L.f4j_dzgemm(M, 1, K, p(A), K, p(xc), 1, p(yc), 1, _native.PREC["auto"], 0))
    assert np.array_equal(yc.view(np.float64), o.mixed_gemm(A, xc.reshape(-1, 1)).reshape(-1).view(np.float64))
    Kc = 37
    Ac, xcc, xr = wide_c(rng, M, Kc), wide_c(rng, Kc), wide(rng, Kc)
    y = zgemv(Ac, xcc)
    assert np.array_equal(y.view(np.float64), o.cplx_gemm(Ac, xcc.reshape(-1, 1)).reshape(-1).view(np.float64))
    yc = np.full(M, np.nan + 0j)
    _native.check(L.f4j_zdgemm(M, 1, Kc, p(Ac), Kc, p(xr), 1, p(yc), 1, _native.PREC["auto"], 0))
    assert np.array_equal(yc.view(np.float64), o.mixed_gemm(Ac, xr.reshape(-1, 1)).reshape(-1).view(np.float64))
