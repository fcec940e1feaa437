"""Pins the CPU oracle against the reference's own golden vectors and invariants,
and against exact arithmetic where no reference vector exists.

T: = src/test/java/org/flag4j/ of the reference; the fixtures were transcribed
from those files by tests/golden/transcribe_fixtures.py.
"""
from fractions import Fraction

import numpy as np
import pytest

from conftest import bits, cplx_mat, real_mat
from oracle import oracle as o

REAL_NN = {
    "standard": o.real_standard,
    "reordered": o.real_reordered,
    "blocked_reordered_bs64": lambda A, B: o.real_blocked_reordered(A, B, 64),
    "blocked_reordered_bs2": lambda A, B: o.real_blocked_reordered(A, B, 2),
    "mt_reordered_T1": lambda A, B: o.real_mt_reordered(A, B, threads=1),
    "mt_reordered_T3": lambda A, B: o.real_mt_reordered(A, B, threads=3),
    "mt_blocked_reordered_T3_bs2": lambda A, B: o.real_mt_blocked_reordered(A, B, threads=3, bs=2),
}


@pytest.mark.parametrize("case", ["real_square", "real_rect"])
@pytest.mark.parametrize("kernel", sorted(REAL_NN))
def test_real_kernels_bitwise_golden(golden, case, kernel):
    """T:linalg/ops/dense/real/MatrixMultiplyTests.java:14-107 (assertArrayEquals is bitwise)."""
    g = golden[case]
    A, B, exp = real_mat(g["a"]), real_mat(g["b"]), real_mat(g["c"])
    act = REAL_NN[kernel](A, B).reshape(-1)
    assert np.array_equal(bits(act), bits(exp))


def test_golden_vectors_reject_fma(golden):
    """A fused multiply-add accumulation is a plausible mistake; the rectangle
    vector must catch it (6 of 24 entries differ)."""
    g = golden["real_rect"]
    A, B, exp = real_mat(g["a"]), real_mat(g["b"]), real_mat(g["c"])
    fma = o.real_reordered_fma(A, B).reshape(-1)
    assert (bits(fma) != bits(exp)).sum() == 6


def test_golden_vectors_reject_transposed_operand(golden):
    g = golden["real_square"]
    A, B, exp = real_mat(g["a"]), real_mat(g["b"]), real_mat(g["c"])
    assert not np.array_equal(o.real_reordered(A, B.T.copy()).reshape(-1), exp)
    assert not np.array_equal(o.real_reordered(B, A).reshape(-1), exp)


def test_real_column_vector_golden(golden):
    """T:linalg/ops/dense/real/MatrixMultiplyTests.java:111-135 (N = 1 kernels)."""
    g = golden["real_colvec"]
    A, x, exp = real_mat(g["a"]), real_mat(g["b"]).reshape(-1), real_mat(g["c"])
    assert np.array_equal(bits(o.real_blocked_vector(A, x, 64)), bits(exp))
    assert np.array_equal(bits(o.real_standard(A, x.reshape(-1, 1)).reshape(-1)), bits(exp))


def blocked_vector_cases():
    """Rows whose results under the N = 1 blocked rule (RealDenseMatMult.java:458-491:
    each 64-wide k block summed from +0.0, the block sums added in k order into
    dest = +0.0) are fixed by hand in exact binary arithmetic.  u = 2^-53 is half
    an ulp of 1, so 1 + u rounds (to even) back to 1, while a sum of tiny terms
    alone is exact.  Each case: (row a, vector x, blocked result, k-ordered result)."""
    u = 2.0**-53
    cases = []
    # K = 128: block 0 = 1 (its 63 tiny terms vanish one by one), block 1 = 64u
    # = 2^-47 exactly, so the blocked result is 1 + 2^-47; one k-ordered chain
    # loses every tiny term and gives 1.  A block size of 32 would give
    # ((1 + 2^-48) + 2^-48) + 2^-48 = 1 + 3 * 2^-48, of 128 the chain's 1.
    cases.append(([1.0] + [u] * 127, [1.0] * 128, 1.0 + 2.0**-47, 1.0))
    # K = 100: the ragged last block holds k = 64..99, 36 terms: 1 + 36u = 1 + 9 * 2^-51.
    cases.append(([1.0] + [u] * 99, [1.0] * 100, 1.0 + 9 * 2.0**-51, 1.0))
    # The tiny terms as products: a = 2^-27, x = 2^-26 (each product is exactly u).
    cases.append(([1.0] + [2.0**-27] * 127, [1.0] + [2.0**-26] * 127, 1.0 + 2.0**-47, 1.0))
    # K = 192, block sums 1, -1, 2^-60 added in k order: (1 - 1) + 2^-60 = 2^-60;
    # in reverse order (2^-60 - 1) + 1 = 0.  The chain gives 2^-60 too.
    a = [0.0] * 192
    a[0], a[64], a[128] = 1.0, -1.0, 2.0**-60
    cases.append((a, [1.0] * 192, 2.0**-60, 2.0**-60))
    # K = 130: tiny terms ahead of the 1 in the last block: block 0 = 64u,
    # block 1 = 64u, block 2 = 1 + u + u... (k = 128: 1, k = 129: u): 1;
    # total ((0 + 2^-47) + 2^-47) + 1 = 1 + 2^-46; the chain: 128u = 2^-46, + 1, + u -> 1 + 2^-46.
    cases.append(([u] * 128 + [1.0, u], [1.0] * 130, 1.0 + 2.0**-46, 1.0 + 2.0**-46))
    # Signed zero: every product -0.0; sums start from +0.0 and +0.0 + -0.0 = +0.0.
    cases.append(([-0.0] * 70, [1.0] * 70, 0.0, 0.0))
    return cases


def test_blocked_vector_exact_cases():
    """The N = 1 blocked kernel against results fixed by hand in exact arithmetic
    (not a restatement of its loop): the 64-wide blocks, the ragged last block,
    block order and the +0.0 starts each decide one case."""
    for a, x, blocked, chain in blocked_vector_cases():
        A = np.array([a], dtype=np.float64)
        xv = np.array(x, dtype=np.float64)
        y = o.real_blocked_vector(A, xv, 64)
        assert bits(y)[0] == bits(np.array([blocked]))[0], (len(a), y[0], blocked)
        ys = o.real_standard(A, xv.reshape(-1, 1)).reshape(-1)
        assert bits(ys)[0] == bits(np.array([chain]))[0], (len(a), ys[0], chain)
    # a plausible mistake, a block size of 32, fails the first case
    a, x, blocked, _ = blocked_vector_cases()[0]
    y32 = o.real_blocked_vector(np.array([a]), np.array(x), 32)
    assert y32[0] == 1.0 + 3 * 2.0**-48 and y32[0] != blocked
    # rows are independent (the 64-row blocks of the i loop change nothing):
    # 130 copies of the first case, negated on odd rows, in one call
    A = np.array([a] * 130) * np.where(np.arange(130) % 2, -1.0, 1.0)[:, None]
    y = o.real_blocked_vector(A, np.array(x), 64)
    assert np.array_equal(y, np.where(np.arange(130) % 2, -blocked, blocked))


def test_real_api_mult_and_pow(golden):
    """T:arrays/dense/matrix/MatrixMultTests.java:26-47 and :193-225 (pow(3) = A*(A*A),
    Matrix.java:1932-1950)."""
    g = golden["real_api_mult"]
    assert np.array_equal(bits(o.real_reordered(real_mat(g["a"]), real_mat(g["b"]))), bits(real_mat(g["c"])))
    g = golden["real_api_pow"]
    A = real_mat(g["a"])
    I = np.eye(3)
    p3 = o.real_reordered(o.real_reordered(I, A), o.real_reordered(A, A))
    assert np.array_equal(bits(p3), bits(real_mat(g["c3"])))
    assert np.array_equal(real_mat(g["c0"]), I)
    assert np.array_equal(real_mat(g["c1"]), A)


@pytest.mark.parametrize("case", ["real_T_square", "real_T_rect", "real_api_multT"])
def test_real_mult_transpose_invariant(golden, case):
    """multTranspose(B) == mult(B.T()) exactly: RealDenseMatMultTransposeTests.java:13-69,
    MatrixMultTests.java:250-266.  The fixture's B literal is the K x N factor;
    multTranspose receives its transpose (N x K)."""
    g = golden[case]
    A, B = real_mat(g["a"]), real_mat(g["b"])
    nn = o.real_reordered(A, B)
    nt = o.real_mult_transpose(A, B.T.copy())
    assert np.array_equal(bits(nn), bits(nt))


@pytest.mark.parametrize("case", ["cplx_square", "cplx_rect", "cplx_api_mult"])
def test_complex_bitwise_golden(golden, case):
    """T:linalg/ops/dense/complex/CMatrixMultiplyTests.java:17-131,
    T:arrays/dense/complex_matrix/CMatrixMultTests.java:64-87."""
    g = golden[case]
    A, B, exp = cplx_mat(g["a"]), cplx_mat(g["b"]), cplx_mat(g["c"])
    act = o.cplx_gemm(A, B).reshape(exp.shape)
    assert np.array_equal(act.real, exp.real) and np.array_equal(act.imag, exp.imag)
    mt = o.cplx_mt_reordered(A, B, threads=2).reshape(exp.shape)
    assert np.array_equal(mt.view(np.float64), act.view(np.float64))


def test_complex_column_vector_golden(golden):
    g = golden["cplx_colvec"]
    A, B, exp = cplx_mat(g["a"]), cplx_mat(g["b"]), cplx_mat(g["c"])
    assert np.array_equal(o.cplx_gemm(A, B).reshape(-1), exp)


def test_real_matrix_vector_golden(golden):
    """T:arrays/dense/matrix/MatrixVectorTests.java:27-45 (Matrix.mult(Vector), small path:
    standardVector; with K <= 64 the blocked form gives the same bits)."""
    g = golden["real_api_matvec"]
    A, x, exp = real_mat(g["a"]), real_mat(g["b"]), real_mat(g["c"])
    assert np.array_equal(bits(o.real_standard(A, x.reshape(-1, 1)).reshape(-1)), bits(exp))
    assert np.array_equal(bits(o.real_blocked_vector(A, x, 64)), bits(exp))


def test_complex_matrix_vector_golden(golden):
    """T:arrays/dense/complex_matrix/CMatrixMatVecMultTests.java:62-84 (CMatrix.mult(CVector))."""
    g = golden["cplx_api_matvec"]
    A, B, exp = cplx_mat(g["a"]), cplx_mat(g["b"]), cplx_mat(g["c"])
    act = o.cplx_gemm(A, B).reshape(-1)
    assert np.array_equal(act.view(np.float64), exp.view(np.float64))
    # a swapped product term (re*im + re*im) is a plausible mistake the vector must reject
    wrong = o.cplx_gemm(A.conj(), B).reshape(-1)
    assert not np.array_equal(wrong, exp)


MIXED = ["real_cmat_api_mult", "cplx_rmat_api_mult", "real_cvec_api_matvec", "cplx_vec_api_matvec"]


def _mixed_operands(g, case):
    a = cplx_mat(g["a"]) if case.startswith("cplx") else real_mat(g["a"])
    b = real_mat(g["b"]) if case.startswith("cplx") else cplx_mat(g["b"])
    return a, (b.reshape(-1, 1) if b.ndim == 1 else b), cplx_mat(g["c"])


@pytest.mark.parametrize("case", MIXED)
def test_mixed_golden(golden, case):
    """T:arrays/dense/matrix/MatrixMultTests.java:66-87, complex_matrix/CMatrixMultTests.java:20-41,
    matrix/MatrixVectorTests.java:62-83, complex_matrix/CMatrixMatVecMultTests.java:22-45
    (CVector / CMatrix equality: == per component, M:numbers/Complex128.java:211-219)."""
    A, B, exp = _mixed_operands(golden[case], case)
    act = o.mixed_gemm(A, B).reshape(exp.shape)
    assert np.all((act.real == exp.real) & (act.imag == exp.imag))
    # plausible mistakes: a conjugated complex factor, a transposed real factor
    if np.iscomplexobj(B):
        bad = o.mixed_gemm(A, B.conj()).reshape(exp.shape)
    else:
        bad = o.mixed_gemm(A.conj(), B).reshape(exp.shape)
    assert not np.all((bad.real == exp.real) & (bad.imag == exp.imag))


def test_mixed_equals_complex_on_promoted_factor():
    """With no zero entries (so no signed-zero differences) the mixed order equals the
    complex kernel's on the real factor promoted to complex: ar*br - 0*bi = ar*br exactly."""
    rng = np.random.default_rng(4)
    A = rng.standard_normal((9, 70)) * 10.0 ** rng.integers(-5, 6, (9, 70))
    B = (rng.standard_normal((70, 5)) + 1j * rng.standard_normal((70, 5))) * 10.0 ** rng.integers(-5, 6, (70, 5))
    assert np.array_equal(o.mixed_gemm(A, B).view(np.float64), o.cplx_gemm(A.astype(complex), B).view(np.float64))
    Bt = B.T.copy()
    At = A.T.copy()
    assert np.array_equal(o.mixed_gemm(Bt, At).view(np.float64), o.cplx_gemm(Bt, At.astype(complex)).view(np.float64))


def test_matrix_vector_shape_errors(golden):
    """Size mismatches raise LinearAlgebraException before any device work
    (MatrixVectorTests.java:47-58, CMatrixMatVecMultTests.java:86-97)."""
    import flag4j_amd as f4j
    g = golden["real_api_matvec"]
    with pytest.raises(f4j.LinearAlgebraException):
        f4j.Matrix(real_mat(g["a"])).mult(f4j.Vector(real_mat(g["b_bad"])))
    g = golden["cplx_api_matvec"]
    with pytest.raises(f4j.LinearAlgebraException):
        f4j.CMatrix(cplx_mat(g["a"])).mult(f4j.CVector(cplx_mat(g["b_bad"]).reshape(-1)))
    g = golden["real_cmat_api_mult"]                # MatrixMultTests.java:89-102
    with pytest.raises(f4j.LinearAlgebraException):
        f4j.Matrix(real_mat(g["a"])).mult(f4j.CMatrix(cplx_mat(g["b_bad"])))
    g = golden["cplx_rmat_api_mult"]                # CMatrixMultTests.java:43-62
    with pytest.raises(f4j.LinearAlgebraException):
        f4j.CMatrix(cplx_mat(g["a"])).mult(f4j.Matrix(real_mat(g["b_bad"])))
    g = golden["real_cvec_api_matvec"]              # MatrixVectorTests.java:86-98
    with pytest.raises(f4j.LinearAlgebraException):
        f4j.Matrix(real_mat(g["a"])).mult(f4j.CVector(cplx_mat(g["b_bad"])))
    g = golden["cplx_vec_api_matvec"]               # CMatrixMatVecMultTests.java:47-59
    with pytest.raises(f4j.LinearAlgebraException):
        f4j.CMatrix(cplx_mat(g["a"])).mult(f4j.Vector(real_mat(g["b_bad"]).reshape(-1)))


def test_complex_pow_golden(golden):
    """T:arrays/dense/complex_matrix/CMatrixMultTests.java:198-219."""
    g = golden["cplx_api_pow"]
    A = cplx_mat(g["a"])
    I = np.eye(3, dtype=np.complex128)
    p3 = o.cplx_gemm(o.cplx_gemm(I, A), o.cplx_gemm(A, A))
    assert np.array_equal(p3, cplx_mat(g["c3"]))


def test_complex_golden_rejects_swapped_product_terms(golden):
    """A plausible mistake, ar*bi - ai*br in the real part, must fail the vector."""
    g = golden["cplx_square"]
    A, B, exp = cplx_mat(g["a"]), cplx_mat(g["b"]), cplx_mat(g["c"])
    wrong = (A.real @ B.real + A.imag @ B.imag) + 1j * (A.real @ B.imag - A.imag @ B.real)
    assert not np.allclose(wrong.reshape(-1), exp)


@pytest.mark.parametrize("case", ["cplx_T_square", "cplx_T_rect"])
def test_complex_mult_transpose_invariant(golden, case):
    """ComplexDenseMatMultTransposeTests.java:40-107: multTranspose(B) == mult(B.T())."""
    g = golden[case]
    A, B = cplx_mat(g["a"]), cplx_mat(g["b"])
    nn = o.cplx_gemm(A, B)
    nt = o.cplx_gemm(A, B.T.copy(), "N", "T")
    assert np.array_equal(nn.view(np.float64), nt.view(np.float64))


def test_complex_conjugate_transpose_operand():
    """op(A) = A^H read on the fly equals multiplying the materialised A.H()
    (conjugation is exact, Complex128.java:513-516)."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((7, 5)) + 1j * rng.standard_normal((7, 5))
    B = rng.standard_normal((7, 4)) + 1j * rng.standard_normal((7, 4))
    fused = o.cplx_gemm(A, B, "C", "N")
    mat = o.cplx_gemm(A.conj().T.copy(), B)
    assert np.array_equal(fused.view(np.float64), mat.view(np.float64))


def test_integer_inputs_exact():
    """Integer-valued inputs have exact products and sums: every kernel must give
    the exact integer product."""
    rng = np.random.default_rng(11)
    A = rng.integers(-50, 50, (33, 47)).astype(np.float64)
    B = rng.integers(-50, 50, (47, 29)).astype(np.float64)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    for f in REAL_NN.values():
        assert np.array_equal(f(A, B), exact)
    assert np.array_equal(o.real_mult_transpose(A, B.T.copy()), exact)


def test_dot2_against_exact_rationals():
    """The accuracy oracle (compensated Dot2) against exact rational dot products."""
    rng = np.random.default_rng(7)
    M, N, K = 6, 5, 300
    A = rng.standard_normal((M, K)) * 10.0 ** rng.uniform(-8, 8, (M, K))
    B = rng.standard_normal((K, N)) * 10.0 ** rng.uniform(-8, 8, (K, N))
    I, J = np.repeat(np.arange(M), N), np.tile(np.arange(N), M)
    hi, lo, ab = o.real_dot2(A, B, I, J)
    for s in range(len(I)):
        ex = sum(Fraction(A[I[s], k]) * Fraction(B[k, J[s]]) for k in range(K))
        err = abs(Fraction(hi[s]) + Fraction(lo[s]) - ex)
        assert float(err) <= 2.0**-100 * ab[s] + 2.0**-1074
    # transposed B
    hi2, lo2, _ = o.real_dot2(A, B.T.copy(), I, J, transB=True)
    assert np.array_equal(hi, hi2) and np.array_equal(lo, lo2)


def test_cplx_dot2_against_exact_rationals():
    rng = np.random.default_rng(8)
    M, N, K = 4, 3, 50
    A = rng.standard_normal((K, M)) + 1j * rng.standard_normal((K, M))
    B = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
    I, J = np.repeat(np.arange(M), N), np.tile(np.arange(N), M)
    rh, rl, ih, il, ab = o.cplx_dot2(A, B, I, J, "C", "N")
    for s in range(len(I)):
        i, j = I[s], J[s]
        er = sum(Fraction(A[k, i].real) * Fraction(B[k, j].real) + Fraction(A[k, i].imag) * Fraction(B[k, j].imag)
                 for k in range(K))
        ei = sum(Fraction(A[k, i].real) * Fraction(B[k, j].imag) - Fraction(A[k, i].imag) * Fraction(B[k, j].real)
                 for k in range(K))
        assert float(abs(Fraction(rh[s]) + Fraction(rl[s]) - er)) <= 2.0**-100 * ab[s]
        assert float(abs(Fraction(ih[s]) + Fraction(il[s]) - ei)) <= 2.0**-100 * ab[s]


def test_transpose_selector_reference_table():
    """T:linalg/ops/RealDenseMatrixMultiplyDispatcherTests.java:56-64 (rows list :19)."""
    rows = [5, 25, 50, 150, 250, 500, 1024, 2048, 4096]
    exp = ["MULT_T", "MULT_T", "MULT_T_BLOCKED", "MULT_T_CONCURRENT", "MULT_T_CONCURRENT", "MULT_T_CONCURRENT",
           "MULT_T_CONCURRENT", "MULT_T_BLOCKED_CONCURRENT", "MULT_T_BLOCKED_CONCURRENT"]
    assert [o.select_transpose(r) for r in rows] == exp


def test_real_dispatch_selections_for_the_workloads():
    """ReDeMatMultDispatcher.getFunc (ReDeMatMultDispatcher.java:183-237) on the
    benchmark shapes, as derived by hand from the source: (m, n=inner, k=B.cols)."""
    assert o.select_real(512, 512, 512) == "MT_REORD"
    assert o.select_real(8192, 8192, 8192) == "MT_BLK_REORD"
    assert o.select_real(16384, 4096, 16384) == "MT_REORD"
    assert o.select_real(4, 3, 2) == "SMALL_STANDARD"
    assert o.select_real(40, 40, 40) == "MT_STRD"
    assert o.select_real(1000, 5, 1000) == "STRD"       # m == maxDim, minDim < 10
    assert o.select_real(20, 1000, 600) == "MT_STRD"    # n == maxDim, m < 25
    assert o.select_real(30, 600, 1000) == "MT_REORD"   # k == maxDim
    assert o.select_real(300, 300, 1) == "MT_BLK_VEC"
    assert o.select_real(100, 100, 1) == "BLK_VEC"
    assert o.select_cplx(300, 300, 1) == "MT_STRD_VEC"
    assert o.select_cplx(2048, 1536, 2048) == "MT_REORD"


# (m, n = inner size, k = B.cols) -> kernel, on both sides of every threshold of
# ReDeMatMultDispatcher.dispatch/getFunc (ReDeMatMultDispatcher.java:157-237, defaults of
# R:configs/ReDeMatMult.properties: smallThreshold 2450, aspectThreshold 2.0,
# squareMtStandardThreshold 50, squareMtReorderedThreshold 2048, minDimSmallThreshold 10,
# wideMtReorderedThreshold 25, squareStandardVecThreshold 256), read off the source text.
# Each pair differs in its result, so a "<" written as "<=" (or a branch taken in
# another order) at any threshold flips one entry.
REAL_BOUNDARIES = [
    ((40, 31, 39), "SMALL_STANDARD"),   # 31 * 40 + 31 * 39 = 2449 < 2450
    ((35, 35, 35), "MT_STRD"),          # 2 * 35^2 = 2450: not small; square, n = 35 < 50
    ((255, 255, 1), "BLK_VEC"),         # vector, aspect 1 <= 8, m = 255 < 256
    ((256, 256, 1), "MT_BLK_VEC"),      # m = 256
    ((2048, 256, 1), "MT_BLK_VEC"),     # aspect exactly 8.0 = 4 * aspectThreshold
    ((2049, 256, 1), "BLK_VEC"),        # aspect 8.004 > 8
    ((256, 2048, 1), "MT_BLK_VEC"),     # aspect 8.0 with n the larger
    ((256, 2049, 1), "BLK_VEC"),
    ((49, 49, 49), "MT_STRD"),          # square, n = 49 < 50
    ((50, 50, 50), "MT_REORD"),         # n = 50
    ((2047, 2047, 2047), "MT_REORD"),   # n = 2047 < 2048
    ((2048, 2048, 2048), "MT_BLK_REORD"),
    ((98, 49, 98), "MT_STRD"),          # aspect exactly 2.0: still "square"
    ((99, 49, 98), "MT_REORD"),         # aspect 2.02: m == maxDim, minDim 49 >= 10
    ((4096, 2048, 4096), "MT_BLK_REORD"),
    ((4097, 2048, 4096), "MT_REORD"),
    ((1000, 9, 1000), "STRD"),          # m == maxDim (ties with k: m is tested first), minDim 9 < 10
    ((1000, 10, 1000), "MT_REORD"),
    ((24, 1000, 600), "MT_STRD"),       # n == maxDim, m = 24 < 25
    ((25, 1000, 600), "MT_REORD"),
    ((9, 1000, 600), "STRD"),           # n == maxDim, minDim 9 < 10
    ((10, 1000, 600), "MT_STRD"),       # minDim 10, m = 10 < 25
    ((5, 1000, 1000), "STRD"),          # n and k tie for maxDim: the n branch is tested first
    ((9, 600, 1000), "REORD"),          # k == maxDim, minDim 9
    ((10, 600, 1000), "MT_REORD"),
]

# Cm128DeMatMultDispatcher.dispatch/getFunc (Cm128DeMatMultDispatcher.java:155-226): the same
# tree for N > 1; every N = 1 product (above the small path) is MT_STRD_VEC, with no
# squareStandardVecThreshold.
CPLX_BOUNDARIES = [
    ((40, 31, 39), "SMALL_STANDARD"),
    ((35, 35, 35), "MT_STRD"),
    ((255, 255, 1), "MT_STRD_VEC"),
    ((2049, 256, 1), "MT_STRD_VEC"),
    ((1224, 1, 1), "SMALL_STANDARD"),   # 1224 + 1 < 2450
    ((49, 49, 49), "MT_STRD"),
    ((50, 50, 50), "MT_REORD"),
    ((2047, 2047, 2047), "MT_REORD"),
    ((2048, 2048, 2048), "MT_BLK_REORD"),
    ((98, 49, 98), "MT_STRD"),
    ((99, 49, 98), "MT_REORD"),
    ((1000, 9, 1000), "STRD"),
    ((24, 1000, 600), "MT_STRD"),
    ((25, 1000, 600), "MT_REORD"),
    ((5, 1000, 1000), "STRD"),
    ((9, 600, 1000), "REORD"),
    ((10, 600, 1000), "MT_REORD"),
]


@pytest.mark.parametrize("shape,kernel", REAL_BOUNDARIES)
def test_real_selector_threshold_boundaries(shape, kernel):
    assert o.select_real(*shape) == kernel


@pytest.mark.parametrize("shape,kernel", CPLX_BOUNDARIES)
def test_cplx_selector_threshold_boundaries(shape, kernel):
    assert o.select_cplx(*shape) == kernel


def test_selector_boundaries_follow_configured_thresholds():
    """The thresholds come from the configuration, not constants: moving one moves
    its boundary (the reference reads them from R:configs/*.properties,
    ReDeMatMultDispatchConfigs.java:52-117)."""
    assert o.select_real(35, 35, 35, small_threshold=2451) == "SMALL_STANDARD"
    assert o.select_real(49, 49, 49, sq_mt_standard=49) == "MT_REORD"
    assert o.select_real(99, 49, 98, aspect_threshold=2.1) == "MT_STRD"
    assert o.select_real(255, 255, 1, sq_seq_vec=255) == "MT_BLK_VEC"
    assert o.select_real(10, 600, 1000, min_dim_small=11) == "REORD"
    assert o.select_real(25, 1000, 600, wide_mt_reordered=26) == "MT_STRD"
    assert o.select_real(2047, 2047, 2047, sq_mt_reordered=2047) == "MT_BLK_REORD"


def test_java_random_known_answers():
    """Known outputs of java.util.Random (Java SE specification of the LCG)."""
    assert o.JavaRandom(42).next_int() == -1170105035
    assert o.JavaRandom(0).next_double() == 0.730967787376657


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_real_entries_equal_the_pinned_kernels(golden, opA, opB):
    """Sampled entries in the reference's arithmetic equal the full products of the
    golden-pinned kernels bit for bit (all op pairs), and the golden vectors themselves."""
    rng = np.random.default_rng(11)
    A = rng.standard_normal((37, 53)) * 10.0 ** rng.integers(-6, 7, (37, 53))
    B = rng.standard_normal((53, 29)) * 10.0 ** rng.integers(-6, 7, (53, 29))
    full = o.real_reordered(A, B)
    Ast = A if opA == "N" else A.T.copy()
    Bst = B if opB == "N" else B.T.copy()
    I, J = np.repeat(np.arange(37), 29), np.tile(np.arange(29), 37)
    got = o.real_entries(Ast, Bst, I, J, opA, opB)
    assert np.array_equal(bits(got), bits(full[I, J]))
    g = golden["real_rect"]
    Ag, Bg, exp = real_mat(g["a"]), real_mat(g["b"]), real_mat(g["c"])
    M, N = Ag.shape[0], Bg.shape[1]
    I, J = np.repeat(np.arange(M), N), np.tile(np.arange(N), M)
    assert np.array_equal(bits(o.real_entries(Ag, Bg, I, J)), bits(exp))


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("C", "N"), ("N", "T"), ("T", "C")])
def test_cplx_entries_equal_the_pinned_kernel(golden, opA, opB):
    rng = np.random.default_rng(12)
    A = rng.standard_normal((21, 17)) + 1j * rng.standard_normal((21, 17))
    B = rng.standard_normal((17, 19)) + 1j * rng.standard_normal((17, 19))
    Ast = A if opA == "N" else A.T.copy()
    Bst = B if opB == "N" else B.T.copy()
    full = o.cplx_gemm(Ast, Bst, opA, opB)
    M, N = full.shape
    I, J = np.repeat(np.arange(M), N), np.tile(np.arange(N), M)
    got = o.cplx_entries(Ast, Bst, I, J, opA, opB)
    assert np.array_equal(got.view(np.float64), full[I, J].view(np.float64))
    g = golden["cplx_square"]
    Ag, Bg, exp = cplx_mat(g["a"]), cplx_mat(g["b"]), cplx_mat(g["c"])
    M, N = Ag.shape[0], Bg.shape[1]
    I, J = np.repeat(np.arange(M), N), np.tile(np.arange(N), M)
    assert np.array_equal(o.cplx_entries(Ag, Bg, I, J).view(np.float64), exp.reshape(-1).view(np.float64))


def test_concurrent_standard_kernels(golden):
    """concurrentStandard (RealDenseMatMult.java:243-272, DenseSemiringMatMult.java:247-287)
    gives the golden vectors' bits and the pinned sequential kernels' bits for any thread count."""
    g = golden["real_rect"]
    A, B, exp = real_mat(g["a"]), real_mat(g["b"]), real_mat(g["c"])
    for T in (1, 2, 5):
        assert np.array_equal(bits(o.real_mt_standard(A, B, threads=T).reshape(-1)), bits(exp))
    g = golden["cplx_rect"]
    A, B, exp = cplx_mat(g["a"]), cplx_mat(g["b"]), cplx_mat(g["c"])
    for T in (1, 3):
        got = o.cplx_mt_standard(A, B, threads=T).reshape(exp.shape)
        assert np.array_equal(got.view(np.float64), exp.view(np.float64))
    rng = np.random.default_rng(13)
    A = rng.standard_normal((40, 70)) * 10.0 ** rng.integers(-5, 6, (40, 70))
    B = rng.standard_normal((70, 33)) * 10.0 ** rng.integers(-5, 6, (70, 33))
    assert np.array_equal(bits(o.real_mt_standard(A, B, threads=4)), bits(o.real_standard(A, B)))
    Az = A[:, :35] + 1j * A[:, 35:]
    Bz = B[:35] + 1j * B[35:]
    assert np.array_equal(o.cplx_mt_standard(Az, Bz, threads=3).view(np.float64),
                          o.cplx_gemm(Az, Bz).view(np.float64))


@pytest.mark.parametrize("case", ["lstsq_vec_1", "lstsq_vec_2", "lstsq_mat_1", "lstsq_mat_2"])
def test_lstsq_golden(golden, case):
    """T:linalg/solvers/RealLstsqSolverTests.java:17-98 (assertEquals on Vector / Matrix is
    Arrays.equals: bitwise)."""
    g = golden[case]
    A, B, exp = real_mat(g["a"]), real_mat(g["b"]), real_mat(g["x"])
    st, X = o.lstsq(A, B)
    assert st == 0
    assert np.array_equal(bits(X.reshape(-1)), bits(exp.reshape(-1))), (X, exp)


def test_lstsq_properties():
    """Q from the restated QR is orthonormal and Q R = A (to rounding); the least-squares
    residual is orthogonal to the columns of A; a rank-deficient A is reported singular,
    a wide A (M < N) rejected."""
    rng = np.random.default_rng(21)
    A = rng.standard_normal((40, 12))
    T, qf, Q = o.qr(A)
    Rm = np.triu(T[:12])
    assert np.allclose(Q.T @ Q, np.eye(12), atol=1e-14)
    assert np.allclose(Q @ Rm, A, atol=1e-13)
    B = rng.standard_normal((40, 3))
    st, X = o.lstsq(A, B)
    assert st == 0
    assert np.allclose(A.T @ (A @ X - B), 0, atol=1e-12)
    assert np.allclose(X, np.linalg.lstsq(A, B, rcond=None)[0], atol=1e-12)
    A2 = A.copy()
    A2[:, 5] = 0.0                                        # R[5][5] = 0: det R = 0
    assert o.lstsq(A2, B)[0] == 1
    A3 = A.copy()
    A3[:, 5] = A3[:, 4]                                   # R[5][5] ~ u only: det R is far above eps n,
    assert o.lstsq(A3, B)[0] == 0                         # so the back solver does not flag it
    assert o.lstsq(A[:10], B[:10])[0] == 2
