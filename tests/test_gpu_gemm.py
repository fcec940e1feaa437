"""GPU parity of the dense GEMM path against the oracle (the C restatement of
Flag4j's kernels), through the C ABI.

Bars:
  * STRICT precision (and the dispatcher's small path): bit-for-bit equal to
    the reference's golden vectors and to the oracle.
  * FP64 MFMA precision: componentwise |C_gpu - C_flag4j|_ij <= 2 K u (|A||B|)_ij
    with u = 2^-53 (both results are within K u (|A||B|) of the exact product);
    for complex the bound uses K~ = 2K and |z| = |re| + |im|.
"""
import concurrent.futures as cf

import numpy as np
import pytest

import flag4j_amd as f4j
from flag4j_amd import device, rng
from conftest import bits, cplx_mat, real_mat
from oracle import oracle as o

pytestmark = pytest.mark.gpu
U = 2.0**-53


def real_bound_ok(C, ref, A, B, transB=False, c=2.0):
    absab = np.abs(A) @ (np.abs(B).T if transB else np.abs(B))
    K = A.shape[1]
    err = np.abs(C - ref)
    ok = err <= c * K * U * absab
    return bool(ok.all()), float(np.max(err / np.maximum(absab, 1e-300)))


def cplx_bound_ok(C, ref, Aop, Bop, c=2.0):
    """Aop, Bop are op(A), op(B) as dense complex arrays."""
    nA = np.abs(Aop.real) + np.abs(Aop.imag)
    nB = np.abs(Bop.real) + np.abs(Bop.imag)
    absab = nA @ nB
    K2 = 2 * Aop.shape[1]
    er = np.abs(C.real - ref.real)
    ei = np.abs(C.imag - ref.imag)
    lim = c * K2 * U * absab
    return bool((er <= lim).all() and (ei <= lim).all())


# ---------------------------------------------------------------- golden vectors

@pytest.mark.parametrize("case", ["real_square", "real_rect", "real_api_mult"])
def test_golden_real_mult_bitwise(golden, case):
    """The small path (ReDeMatMultDispatcher.java:159-168) runs the strict kernel:
    results equal the reference's vectors bit for bit."""
    g = golden[case]
    A, B, exp = real_mat(g["a"]), real_mat(g["b"]), real_mat(g["c"]).reshape(-1)
    C = f4j.Matrix(A).mult(f4j.Matrix(B))
    assert np.array_equal(bits(C.data), bits(exp))


def test_golden_real_pow_bitwise(golden):
    g = golden["real_api_pow"]
    A = f4j.Matrix(real_mat(g["a"]))
    assert A.pow(3).equals(f4j.Matrix(real_mat(g["c3"])))
    assert np.array_equal(bits(A.pow(3).data), bits(real_mat(g["c3"]).reshape(-1)))
    assert A.pow(0).equals(f4j.Matrix(real_mat(g["c0"])))
    assert A.pow(1).equals(f4j.Matrix(real_mat(g["c1"])))


@pytest.mark.parametrize("case", ["real_T_square", "real_T_rect", "real_api_multT"])
def test_golden_real_mult_transpose_invariant(golden, case):
    """MatrixMultTests.java:250-266: A.multTranspose(B) equals A.mult(B.T())."""
    g = golden[case]
    A = f4j.Matrix(real_mat(g["a"]))
    Bt = f4j.Matrix(real_mat(g["b"])).T()          # stored N x K
    assert A.multTranspose(Bt).equals(A.mult(Bt.T()))
    assert np.array_equal(bits(A.multTranspose(Bt).data),
                          bits(o.real_reordered(A.numpy(), real_mat(g["b"])).reshape(-1)))


@pytest.mark.parametrize("case", ["cplx_square", "cplx_rect", "cplx_api_mult", "cplx_colvec"])
def test_golden_complex_mult_bitwise(golden, case):
    g = golden[case]
    A, B, exp = cplx_mat(g["a"]), cplx_mat(g["b"]), cplx_mat(g["c"]).reshape(-1)
    C = f4j.CMatrix(A).mult(f4j.CMatrix(B))
    assert np.array_equal(C.data.view(np.float64), exp.view(np.float64)) or C.equals(f4j.CMatrix(C.shape, exp))
    assert C.equals(f4j.CMatrix(C.shape, exp))


def test_golden_complex_pow(golden):
    g = golden["cplx_api_pow"]
    A = f4j.CMatrix(cplx_mat(g["a"]))
    assert A.pow(3).equals(f4j.CMatrix(cplx_mat(g["c3"])))
    assert A.pow(0).equals(f4j.CMatrix(cplx_mat(g["c0"])))


@pytest.mark.parametrize("case", ["cplx_T_square", "cplx_T_rect"])
def test_golden_complex_mult_transpose_invariant(golden, case):
    g = golden[case]
    A = f4j.CMatrix(cplx_mat(g["a"]))
    Bt = f4j.CMatrix(cplx_mat(g["b"])).T()
    assert A.multTranspose(Bt).equals(A.mult(Bt.T()))


# ---------------------------------------------------------------- strict kernel

@pytest.mark.parametrize("opB", ["N", "T"])
def test_strict_real_matches_oracle_bitwise(opB):
    A = rng.uniform(11, 100, 90)
    B = rng.uniform(12, 90, 70)
    C = np.empty((100, 70))
    Bs = B if opB == "N" else B.T.copy()
    ld = 70 if opB == "N" else 90
    f4j._native.check(f4j._native.load().f4j_dgemm(0, f4j._native.OP[opB], 100, 70, 90, A.ctypes.data, 90,
                                                   Bs.ctypes.data, ld, C.ctypes.data, 70, 1, 0))
    assert np.array_equal(bits(C), bits(o.real_reordered(A, B)))


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("C", "N"), ("N", "T"), ("T", "C")])
def test_strict_complex_matches_oracle_bitwise(opA, opB):
    M, N, K = 37, 29, 41
    A = rng.uniform_complex(21, *((M, K) if opA == "N" else (K, M)))
    B = rng.uniform_complex(22, *((K, N) if opB == "N" else (N, K)))
    C = np.empty((M, N), dtype=np.complex128)
    L = f4j._native
    L.check(L.load().f4j_zgemm(L.OP[opA], L.OP[opB], M, N, K, A.ctypes.data, A.shape[1], B.ctypes.data, B.shape[1],
                               C.ctypes.data, N, 1, 0, 0))
    assert np.array_equal(C.view(np.float64), o.cplx_gemm(A, B, opA, opB).view(np.float64))


# ---------------------------------------------------------------- FP64 MFMA kernel

def test_config1_real_512_cubed():
    """512^3, entries 2r-1 from java.util.Random(1) / (2)."""
    A, B = rng.uniform(1, 512, 512), rng.uniform(2, 512, 512)
    C = f4j.Matrix(A).mult(f4j.Matrix(B), precision="fp64").numpy()
    ref = o.real_mt_reordered(A, B, threads=8)
    ok, worst = real_bound_ok(C, ref, A, B)
    assert ok, worst
    assert worst < 1e-15


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (1, 7, 3), (17, 33, 65), (64, 64, 64), (129, 130, 131),
                                   (255, 257, 1), (300, 40, 1000), (1000, 3, 17), (2, 1500, 9)])
@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_real_ragged_shapes_all_ops(M, N, K, opA, opB):
    A = rng.uniform(M * 7 + K, *((M, K) if opA == "N" else (K, M)))
    B = rng.uniform(N * 5 + K, *((K, N) if opB == "N" else (N, K)))
    Aop = A if opA == "N" else A.T
    Bop = B if opB == "N" else B.T
    C = np.empty((M, N))
    L = f4j._native
    L.check(L.load().f4j_dgemm(L.OP[opA], L.OP[opB], M, N, K, A.ctypes.data, A.shape[1], B.ctypes.data,
                               B.shape[1], C.ctypes.data, N, 0, 0))
    ref = o.real_reordered(np.ascontiguousarray(Aop), np.ascontiguousarray(Bop))
    assert real_bound_ok(C, ref, Aop, Bop)[0]


def test_real_integer_inputs_exact():
    """Integer inputs: every product and partial sum is exact, so the MFMA result
    must equal the exact integer product."""
    r = np.random.default_rng(3)
    A = r.integers(-64, 64, (200, 300)).astype(np.float64)
    B = r.integers(-64, 64, (300, 150)).astype(np.float64)
    C = f4j.Matrix(A).mult(f4j.Matrix(B), precision="fp64").numpy()
    assert np.array_equal(C, (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64))


def test_real_nn_equals_nt_bitwise():
    """multTranspose(B) == mult(B.T()) exactly also on the MFMA path (both feed the
    same operand values in the same k order)."""
    A, B = rng.uniform(5, 300, 200), rng.uniform(6, 200, 250)
    nn = f4j.Matrix(A).mult(f4j.Matrix(B), precision="fp64")
    nt = f4j.Matrix(A).multTranspose(f4j.Matrix(B.T.copy()), precision="fp64")
    assert np.array_equal(bits(nn.data), bits(nt.data))


def test_k_zero_and_empty():
    L = f4j._native
    C = np.full((3, 4), 7.0)
    A = np.zeros((3, 1))
    B = np.zeros((1, 4))
    L.check(L.load().f4j_dgemm(0, 0, 3, 4, 0, A.ctypes.data, 1, B.ctypes.data, 4, C.ctypes.data, 4, 0, 0))
    assert np.array_equal(C, np.zeros((3, 4)))


def test_leading_dimensions():
    """Sub-matrix views through lda/ldb/ldc larger than the row length."""
    A = rng.uniform(31, 70, 90)
    B = rng.uniform(32, 60, 80)
    C = np.full((50, 100), -1.0)
    L = f4j._native
    L.check(L.load().f4j_dgemm(0, 0, 50, 70, 60, A.ctypes.data, 90, B.ctypes.data, 80, C.ctypes.data, 100, 0, 0))
    ref = o.real_reordered(A[:50, :60].copy(), B[:, :70].copy())
    assert real_bound_ok(C[:, :70], ref, A[:50, :60], B[:, :70])[0]
    assert np.all(C[:, 70:] == -1.0)


def test_deterministic():
    A, B = rng.uniform(41, 640, 320), rng.uniform(42, 320, 704)
    c1 = f4j.Matrix(A).mult(f4j.Matrix(B), precision="fp64").data
    c2 = f4j.Matrix(A).mult(f4j.Matrix(B), precision="fp64").data
    assert np.array_equal(bits(c1), bits(c2))


def test_concurrent_callers():
    """The dispatchers are documented thread-safe (ReDeMatMultDispatcher.java:58-63):
    many host threads calling at once each get their own correct product."""
    shapes = [(100 + 13 * i, 90 + 7 * i, 80 + 11 * i) for i in range(8)]
    ins = [(rng.uniform(100 + i, m, k), rng.uniform(200 + i, k, n)) for i, (m, n, k) in enumerate(shapes)]

    def run(i):
        A, B = ins[i]
        return f4j.Matrix(A).mult(f4j.Matrix(B), precision="fp64").numpy()

    with cf.ThreadPoolExecutor(8) as ex:
        outs = list(ex.map(run, range(8)))
    for (A, B), C in zip(ins, outs):
        assert real_bound_ok(C, o.real_reordered(A, B), A, B)[0]


def test_device_resident_operands():
    M, N, K = 256, 192, 320
    A, B = rng.uniform(51, M, K), rng.uniform(52, K, N)
    with device.DeviceBuffer.from_array(A) as dA, device.DeviceBuffer.from_array(B) as dB, \
            device.DeviceBuffer(M * N * 8) as dC:
        device.dgemm_dev("N", "N", M, N, K, dA.ptr, K, dB.ptr, N, dC.ptr, N)
        device.sync()
        C = dC.download(np.float64, M * N).reshape(M, N)
    assert real_bound_ok(C, o.real_reordered(A, B), A, B)[0]


def test_config3_8192_sampled_against_exact():
    """8192^3 (the benchmark workload): sampled entries against the compensated
    oracle, normalised by (|A||B|)_ij; plus full-row parity with the Flag4j
    restatement on a few rows."""
    n = 8192
    A, B = rng.uniform(1, n, n), rng.uniform(2, n, n)
    C = f4j.Matrix(A).mult(f4j.Matrix(B), precision="fp64").numpy()
    r = np.random.default_rng(0)
    I, J = r.integers(0, n, 2048), r.integers(0, n, 2048)
    hi, lo, ab = o.real_dot2(A, B, I, J)
    err = np.abs((C[I, J] - hi) - lo) / ab
    assert err.max() <= n * U
    for r0 in (0, 8191):
        ref = o.real_reordered(A[r0:r0 + 1], B)
        assert real_bound_ok(C[r0:r0 + 1], ref, A[r0:r0 + 1], B)[0]


# ---------------------------------------------------------------- complex

def test_config2_complex_2048x1536x2048_rows():
    """CMatrix.mult at 2048x1536 * 1536x2048 (the reference API throws here, defect
    D1; the oracle restates the kernels directly).  Full rows 0..63 and 1984..2047
    compared with the componentwise bound."""
    M, K, N = 2048, 1536, 2048
    A, B = rng.uniform_complex(3, M, K), rng.uniform_complex(4, K, N)
    C = f4j.CMatrix(A).mult(f4j.CMatrix(B), precision="fp64").numpy()
    for r0 in (0, 1984):
        ref = o.cplx_mt_reordered(A, B, C=np.zeros((M, N), dtype=np.complex128), rows=(r0, r0 + 64), threads=8)
        assert cplx_bound_ok(C[r0:r0 + 64], ref[r0:r0 + 64], A[r0:r0 + 64], B)


@pytest.mark.parametrize("opA", ["N", "T", "C"])
@pytest.mark.parametrize("opB", ["N", "T", "C"])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (33, 17, 65), (130, 129, 70)])
def test_complex_all_ops(opA, opB, M, N, K):
    A = rng.uniform_complex(M + 3 * K, *((M, K) if opA == "N" else (K, M)))
    B = rng.uniform_complex(N + 5 * K, *((K, N) if opB == "N" else (N, K)))
    C = np.empty((M, N), dtype=np.complex128)
    L = f4j._native
    L.check(L.load().f4j_zgemm(L.OP[opA], L.OP[opB], M, N, K, A.ctypes.data, A.shape[1], B.ctypes.data,
                               B.shape[1], C.ctypes.data, N, 0, 0, 0))
    ref = o.cplx_gemm(A, B, opA, opB)
    op = {"N": lambda X: X, "T": lambda X: X.T, "C": lambda X: X.conj().T}
    assert cplx_bound_ok(C, ref, op[opA](A), op[opB](B))


def test_complex_hermitian_fused_equals_materialised():
    """A.H().mult(B) with op C fused equals multiplying the materialised A^H, bitwise
    (same operand values, same k order on the matrix cores)."""
    A, B = rng.uniform_complex(61, 300, 200), rng.uniform_complex(62, 300, 150)
    fused = f4j.CMatrix(A).H().mult(f4j.CMatrix(B), precision="fp64")
    mat = f4j.CMatrix(A.conj().T.copy()).mult(f4j.CMatrix(B), precision="fp64")
    assert np.array_equal(fused.data.view(np.float64), mat.data.view(np.float64))


def test_complex_gaussian_integers_exact():
    r = np.random.default_rng(9)
    A = (r.integers(-20, 20, (90, 110)) + 1j * r.integers(-20, 20, (90, 110)))
    B = (r.integers(-20, 20, (110, 70)) + 1j * r.integers(-20, 20, (110, 70)))
    C = f4j.CMatrix(A).mult(f4j.CMatrix(B), precision="fp64").numpy()
    assert np.array_equal(C, A @ B)
