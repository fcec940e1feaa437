"""GPU tests of the split-precision GEMM (FP64 sliced into BF16/FP16 digits).

Accuracy is measured against the compensated oracle (Dot2) on sampled
entries, normalised by (|A||B|)_ij.  Bars stated per test.
"""
import ctypes

import numpy as np
import pytest

import flag4j_amd as f4j
from flag4j_amd import _native as L
from flag4j_amd import rng
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def split_mult(A, B, prec, slices, opA="N", opB="N"):
    M = A.shape[0] if opA == "N" else A.shape[1]
    K = A.shape[1] if opA == "N" else A.shape[0]
    N = B.shape[1] if opB == "N" else B.shape[0]
    C = np.empty((M, N))
    L.check(L.load().f4j_dgemm(L.OP[opA], L.OP[opB], M, N, K, A.ctypes.data, A.shape[1], B.ctypes.data,
                               B.shape[1], C.ctypes.data, N, L.PREC[prec], slices))
    return C, L.last_split_report()


def sampled_err(C, A, B, n=600, seed=0):
    r = np.random.default_rng(seed)
    I, J = r.integers(0, A.shape[0], n), r.integers(0, B.shape[1], n)
    hi, lo, ab = o.real_dot2(A, B, I, J)
    return float(np.max(np.abs((C[I, J] - hi) - lo) / ab))


@pytest.fixture(autouse=True)
def default_split_config():
    L.config_set("splitCheckRows", 4)
    L.config_set("splitTolerance", 1e-15)
    L.config_set("splitFallback", 1)
    yield
    L.config_set("splitCheckRows", 4)
    L.config_set("splitTolerance", 1e-15)
    L.config_set("splitFallback", 1)


@pytest.mark.parametrize("prec", ["split_bf16", "split_fp16"])
def test_small_integers_are_exact(prec):
    """Entries that one digit represents exactly (integers below 2^8 in magnitude
    relative to the row/column maximum): the slice products are exact integer
    GEMMs, so C equals the exact product bit for bit, even with S = 1."""
    r = np.random.default_rng(1)
    A = r.integers(-127, 128, (200, 300)).astype(np.float64)
    B = r.integers(-127, 128, (300, 150)).astype(np.float64)
    A[:, 0] = 127.0                      # pin each row's maximum exponent
    B[0, :] = 127.0
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    for S in (1, 2, 3):
        C, rep = split_mult(A, B, prec, S)
        assert np.array_equal(C, exact), S
        assert rep.slices == S and rep.products == S * (S + 1) // 2


@pytest.mark.parametrize("prec,S", [("split_bf16", 6), ("split_fp16", 5)])
def test_default_slices_reach_fp64_accuracy(prec, S):
    """Uniform[-1,1) inputs at 1024^3: max error over sampled entries <= 1e-15 of
    (|A||B|)_ij, and the on-GPU residual check agrees with the oracle."""
    A, B = rng.uniform(1, 1024, 1024), rng.uniform(2, 1024, 1024)
    C, rep = split_mult(A, B, prec, S)
    err = sampled_err(C, A, B)
    assert err <= 1e-15, err
    assert rep.precision == L.PREC[prec]
    assert rep.checked_rows == 4 and 0 <= rep.max_rel_err <= 1e-15


@pytest.mark.parametrize("prec", ["split_bf16", "split_fp16"])
def test_error_falls_with_slices(prec):
    """Each extra slice adds R bits (9 for BF16, 10 for FP16): the error must fall
    by a large factor per slice until it reaches FP64 rounding."""
    A, B = rng.uniform(3, 512, 768), rng.uniform(4, 768, 512)
    L.config_set("splitCheckRows", 0)
    errs = []
    for S in range(2, 7):
        C, rep = split_mult(A, B, prec, S)
        errs.append(sampled_err(C, A, B))
        assert rep.checked_rows == 0 and rep.max_rel_err == -1.0
    for a, b in zip(errs, errs[1:]):
        assert b <= max(a / 64, 2e-16), errs
    assert errs[0] > 1e-9


@pytest.mark.parametrize("phi,S,bar", [(0, 6, 1e-15), (4, 7, 1e-15), (8, 8, 1e-15)])
def test_wide_dynamic_range(phi, S, bar):
    """Entries (2r-1) 10^(phi (2r'-1)): wider ranges need more slices (the digit
    grid is set by each row's largest entry).  With enough slices the bar holds."""
    A = rng.scaled_uniform(5, 512, 512, phi)
    B = rng.scaled_uniform(6, 512, 512, phi)
    L.config_set("splitCheckRows", 0)
    C, _ = split_mult(A, B, "split_bf16", S)
    assert sampled_err(C, A, B) <= bar


def test_residual_check_fallback_and_error():
    A, B = rng.uniform(7, 256, 256), rng.uniform(8, 256, 256)
    L.config_set("splitTolerance", 1e-30)                 # nothing passes
    C, rep = split_mult(A, B, "split_bf16", 2)
    assert rep.precision == L.PREC["fp64"]                # recomputed in FP64
    ref = o.real_reordered(A, B)
    assert np.all(np.abs(C - ref) <= 2 * 256 * 2.0**-53 * (np.abs(A) @ np.abs(B)))
    L.config_set("splitFallback", 0)
    C = np.empty((256, 256))
    st = L.load().f4j_dgemm(0, 0, 256, 256, 256, A.ctypes.data, 256, B.ctypes.data, 256, C.ctypes.data, 256,
                            L.PREC["split_bf16"], 2)
    assert st == L.ERR_ACCURACY
    assert L.last_split_report().max_rel_err > 1e-12


@pytest.mark.parametrize("opA,opB", [("N", "T"), ("T", "N"), ("T", "T")])
def test_transposed_operands(opA, opB):
    M, N, K = 300, 200, 500
    A = rng.uniform(9, *((M, K) if opA == "N" else (K, M)))
    B = rng.uniform(10, *((K, N) if opB == "N" else (N, K)))
    C, rep = split_mult(A, B, "split_bf16", 6, opA, opB)
    assert rep.precision == L.PREC["split_bf16"]
    Aop = np.ascontiguousarray(A if opA == "N" else A.T)
    Bop = np.ascontiguousarray(B if opB == "N" else B.T)
    assert sampled_err(C, Aop, Bop) <= 1e-15


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (7, 300, 13), (130, 70, 257), (129, 257, 1000)])
@pytest.mark.parametrize("prec", ["split_bf16", "split_fp16"])
def test_ragged_shapes(M, N, K, prec):
    A, B = rng.uniform(M + 11, M, K), rng.uniform(N + 12, K, N)
    C, rep = split_mult(A, B, prec, 6)
    assert rep.precision == L.PREC[prec]
    assert sampled_err(C, A, B, n=200) <= 1e-15


def test_zero_rows_and_columns():
    A, B = rng.uniform(13, 100, 120), rng.uniform(14, 120, 90)
    A[5, :] = 0.0
    B[:, 7] = 0.0
    C, _ = split_mult(A, B, "split_bf16", 6)
    assert np.all(C[5, :] == 0.0) and np.all(C[:, 7] == 0.0)
    mask = np.ones_like(C, dtype=bool)
    mask[5, :] = False
    mask[:, 7] = False
    ref = o.real_reordered(A, B)
    assert np.all(np.abs(C - ref)[mask] <= 2 * 120 * 2.0**-53 * (np.abs(A) @ np.abs(B))[mask])


def test_matrix_api_split():
    A, B = rng.uniform(15, 200, 200), rng.uniform(16, 200, 200)
    C = f4j.Matrix(A).mult(f4j.Matrix(B), precision="split_bf16", slices=6).numpy()
    assert sampled_err(C, A, B) <= 1e-15


# ---------------------------------------------------------------- modular (CRT) form, INT8 matrix cores

def test_crt_integer_inputs_exact():
    """Integer entries below 2^20: with the default moduli (t = 53 bits) every
    scaled entry is exact, the residue products are exact and C is the exact
    integer product."""
    r = np.random.default_rng(2)
    A = r.integers(-(1 << 20), 1 << 20, (150, 260)).astype(np.float64)
    B = r.integers(-(1 << 20), 1 << 20, (260, 170)).astype(np.float64)
    C, rep = split_mult(A, B, "split_int8", 0)
    exact = A.astype(object).dot(B.astype(object))
    assert np.array_equal(C, np.array(exact, dtype=np.float64))
    assert rep.precision == L.PREC["split_int8"] and rep.slices >= 14


def test_crt_default_moduli_fp64_accuracy():
    A, B = rng.uniform(21, 1024, 1024), rng.uniform(22, 1024, 1024)
    C, rep = split_mult(A, B, "split_int8", 0)
    assert sampled_err(C, A, B) <= 1e-15
    assert rep.precision == L.PREC["split_int8"] and 0 <= rep.max_rel_err <= 1e-15


def test_crt_error_falls_with_moduli():
    """Each extra modulus adds ~8 bits to the product range, i.e. ~4 bits to t."""
    A, B = rng.uniform(23, 400, 600), rng.uniform(24, 600, 300)
    L.config_set("splitCheckRows", 0)
    errs = []
    for n in (8, 10, 12, 14, 16):
        C, rep = split_mult(A, B, "split_int8", n)
        assert rep.slices == n
        errs.append(sampled_err(C, A, B))
    for a, b in zip(errs, errs[1:]):
        assert b <= max(a / 100, 2.5e-16), errs
    assert errs[0] > 1e-10


@pytest.mark.parametrize("opA,opB", [("N", "T"), ("T", "N"), ("T", "T")])
def test_crt_transposed(opA, opB):
    M, N, K = 300, 200, 500
    A = rng.uniform(25, *((M, K) if opA == "N" else (K, M)))
    B = rng.uniform(26, *((K, N) if opB == "N" else (N, K)))
    C, rep = split_mult(A, B, "split_int8", 0, opA, opB)
    assert rep.precision == L.PREC["split_int8"]          # delivered without the FP64 fallback
    Aop = np.ascontiguousarray(A if opA == "N" else A.T)
    Bop = np.ascontiguousarray(B if opB == "N" else B.T)
    assert sampled_err(C, Aop, Bop) <= 1e-15


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (7, 300, 13), (130, 70, 257), (129, 257, 1000)])
def test_crt_ragged(M, N, K):
    A, B = rng.uniform(M + 31, M, K), rng.uniform(N + 32, K, N)
    C, rep = split_mult(A, B, "split_int8", 0)
    assert rep.precision == L.PREC["split_int8"]
    assert sampled_err(C, A, B, n=200) <= 1e-15


@pytest.mark.parametrize("N,opB", [(1000, "N"), (2056, "N"), (2056, "T"), (1001, "N")])
def test_crt_column_block_matches_whole(N, opB):
    """Column exponents depend only on the column and t only on K, so the first 512
    columns of a wide product must equal, bit for bit, the product of the first 512
    columns alone; every entry meets the bar (real, and complex 3M on the same shapes)."""
    M, K = 200, 384
    A = rng.uniform(41, M, K)
    B = rng.uniform(42, *((K, N) if opB == "N" else (N, K)))
    C, rep = split_mult(A, B, "split_int8", 0, "N", opB)
    assert rep.precision == L.PREC["split_int8"]
    Bs = np.ascontiguousarray(B[:, :512] if opB == "N" else B[:512, :])
    C512, _ = split_mult(A, Bs, "split_int8", 0, "N", opB)
    assert np.array_equal(C[:, :512].view(np.int64), C512.view(np.int64))
    Bop = np.ascontiguousarray(B if opB == "N" else B.T)
    assert sampled_err(C, A, Bop, n=400) <= 1e-15
    Ac = A + 1j * rng.uniform(43, M, K)
    Bc = Bop + 1j * rng.uniform(44, K, N)
    Cc, rep = zsplit_mult(Ac, Bc, "split_int8", 0, True)
    assert rep.precision == L.PREC["split_int8"]
    assert zsampled_err(Cc, Ac, Bc, "N", "N") <= 1e-15


@pytest.mark.parametrize("phi,n", [(4, 18), (8, 20)])
def test_crt_zero_lines_and_wide_range(phi, n):
    """Entries spread over 10^(+-phi): the default moduli (t ~ 53) leave the small
    entries of a row with too few bits; more moduli raise t (65 and 72 bits at
    K = 256) and restore the bar."""
    A = rng.scaled_uniform(27, 256, 256, phi)
    B = rng.scaled_uniform(28, 256, 256, phi)
    A[3, :] = 0.0
    B[:, 9] = 0.0
    L.config_set("splitCheckRows", 0)
    C, _ = split_mult(A, B, "split_int8", n)
    assert np.all(C[3, :] == 0.0) and np.all(C[:, 9] == 0.0)
    mask = np.ones(C.shape, bool)
    mask[3, :] = False
    mask[:, 9] = False
    I, J = np.nonzero(mask)
    sel = np.random.default_rng(0).choice(len(I), 400, replace=False)
    hi, lo, ab = o.real_dot2(A, B, I[sel], J[sel])
    assert np.max(np.abs((C[I[sel], J[sel]] - hi) - lo) / ab) <= 1e-15


def test_crt_rejects_unrepresentable_k():
    A, B = np.ones((1, 1 << 17)), np.ones((1 << 17, 1))
    C = np.empty((1, 1))
    st = L.load().f4j_dgemm(0, 0, 1, 1, 1 << 17, A.ctypes.data, 1 << 17, B.ctypes.data, 1, C.ctypes.data, 1,
                            L.PREC["split_int8"], 0)
    assert st == L.ERR_INVALID_ARG


def test_crt_moduli_out_of_range_are_rejected():
    """A modulus count outside 2..20 is an invalid argument (as in the batched plans),
    not silently clamped."""
    A, B = rng.uniform(40, 64, 64), rng.uniform(41, 64, 64)
    C = np.empty((64, 64))
    for bad in (1, 21, 25):
        st = L.load().f4j_dgemm(0, 0, 64, 64, 64, A.ctypes.data, 64, B.ctypes.data, 64, C.ctypes.data, 64,
                                L.PREC["split_int8"], bad)
        assert st == L.ERR_INVALID_ARG, bad
        assert b"moduli" in L.load().f4j_last_error()
    C, rep = split_mult(A, B, "split_int8", 20)
    assert rep.slices == 20 and rep.precision == L.PREC["split_int8"]


@pytest.mark.parametrize("prec,S", [("split_bf16", 3), ("split_int8", 0)])
def test_product_timing_report(prec, S):
    """splitTiming = 1 reports the product launches' event time; 0 reports -1."""
    A, B = rng.uniform(29, 512, 512), rng.uniform(30, 512, 512)
    L.config_set("splitTiming", 1)
    try:
        _, rep = split_mult(A, B, prec, S)
        assert 0.0 < rep.product_ms < 1000.0
    finally:
        L.config_set("splitTiming", 0)
    _, rep = split_mult(A, B, prec, S)
    assert rep.product_ms == -1.0


# ---------------------------------------------------------------- complex products (batched engine, batch of one)

def zsplit_mult(A, B, prec, slices, use3m, opA="N", opB="N"):
    M = A.shape[0] if opA == "N" else A.shape[1]
    K = A.shape[1] if opA == "N" else A.shape[0]
    N = B.shape[1] if opB == "N" else B.shape[0]
    A, B = np.ascontiguousarray(A, np.complex128), np.ascontiguousarray(B, np.complex128)
    C = np.empty((M, N), np.complex128)
    L.check(L.load().f4j_zgemm(L.OP[opA], L.OP[opB], M, N, K, A.ctypes.data, A.shape[1], B.ctypes.data,
                               B.shape[1], C.ctypes.data, N, L.PREC[prec], slices, int(use3m)))
    return C, L.last_split_report()


def zsampled_err(C, A, B, opA, opB, n=300, seed=0):
    r = np.random.default_rng(seed)
    I, J = r.integers(0, C.shape[0], n), r.integers(0, C.shape[1], n)
    rh, rl, ih, il, ab = o.cplx_dot2(A, B, I, J, opA, opB)
    return float(np.max(np.maximum(np.abs((C[I, J].real - rh) - rl), np.abs((C[I, J].imag - ih) - il)) / ab))


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("C", "N"), ("N", "T"), ("T", "C")])
@pytest.mark.parametrize("use3m", [True, False])
def test_complex_split_single(opA, opB, use3m):
    M, N, K = 200, 150, 300
    sa = (M, K) if opA == "N" else (K, M)
    sb = (K, N) if opB == "N" else (N, K)
    A = rng.uniform(41, *sa) + 1j * rng.uniform(42, *sa)
    B = rng.uniform(43, *sb) + 1j * rng.uniform(44, *sb)
    C, rep = zsplit_mult(A, B, "split_fp16", 6, use3m, opA, opB)
    assert rep.precision == L.PREC["split_fp16"]
    assert rep.products == (3 if use3m else 4) * 21
    assert 0 <= rep.max_rel_err <= 1e-15
    assert zsampled_err(C, A, B, opA, opB) <= 1e-15


def test_complex_digit_plan_reused_with_new_data():
    """A complex digit-form call keeps its batch plan for the next call with the
    same arguments: refilled device buffers are read afresh, and a call with
    another shape on the same buffers builds a new plan; every result within
    1e-15 of the compensated sums."""
    from flag4j_amd import device
    M, N, K = 160, 120, 200
    cap = 2 * M * K
    with device.DeviceBuffer(cap * 8) as dA, device.DeviceBuffer(cap * 8) as dB, \
            device.DeviceBuffer(2 * M * N * 8) as dC:
        for seed, (m, n, k) in ((71, (M, N, K)), (72, (M, N, K)), (73, (M // 2, N, K // 2))):
            A = rng.uniform(seed, m, k) + 1j * rng.uniform(seed + 100, m, k)
            B = rng.uniform(seed + 200, k, n) + 1j * rng.uniform(seed + 300, k, n)
            dA.upload(np.ascontiguousarray(A).view(np.float64))
            dB.upload(np.ascontiguousarray(B).view(np.float64))
            f4j.zgemm_dev("N", "N", m, n, k, dA.ptr, k, dB.ptr, n, dC.ptr, n, precision="split_bf16", use_3m=True)
            device.sync()
            C = dC.download(np.float64, 2 * m * n).view(np.complex128).reshape(m, n)
            rep = L.last_split_report()
            assert rep.precision == L.PREC["split_bf16"]
            assert zsampled_err(C, A, B, "N", "N") <= 1e-15


def test_complex_split_config2_shape_bf16():
    """CMatrix.mult at 2048 x 1536 x 2048 (uniform[-1,1) parts), split BF16, 3M."""
    A = rng.uniform(45, 2048, 1536) + 1j * rng.uniform(46, 2048, 1536)
    B = rng.uniform(47, 1536, 2048) + 1j * rng.uniform(48, 1536, 2048)
    C, rep = zsplit_mult(A, B, "split_bf16", 6, True)
    assert rep.precision == L.PREC["split_bf16"] and 0 <= rep.max_rel_err <= 1e-15
    assert zsampled_err(C, A, B, "N", "N", n=200) <= 1e-15


def test_complex_split_fallback():
    A = rng.uniform(49, 64, 64) + 1j * rng.uniform(50, 64, 64)
    B = rng.uniform(51, 64, 64) + 1j * rng.uniform(52, 64, 64)
    L.config_set("splitTolerance", 1e-30)
    C, rep = zsplit_mult(A, B, "split_fp16", 2, True)
    assert rep.precision == L.PREC["fp64"]
    assert zsampled_err(C, A, B, "N", "N") <= 1e-15
    L.config_set("splitFallback", 0)
    C = np.empty((64, 64), np.complex128)
    st = L.load().f4j_zgemm(0, 0, 64, 64, 64, A.ctypes.data, 64, B.ctypes.data, 64, C.ctypes.data, 64,
                            L.PREC["split_fp16"], 2, 1)
    assert st == L.ERR_ACCURACY


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("C", "N"), ("N", "T"), ("T", "C")])
@pytest.mark.parametrize("use3m", [True, False])
def test_complex_crt(opA, opB, use3m):
    """Complex products in the modular form: 3 (3M) or 4 real INT8 products, formed from one
    read of each operand per pass and run as one product launch and one reconstruction."""
    M, N, K = 300, 200, 500
    sa = (M, K) if opA == "N" else (K, M)
    sb = (K, N) if opB == "N" else (N, K)
    A = rng.uniform(61, *sa) + 1j * rng.uniform(62, *sa)
    B = rng.uniform(63, *sb) + 1j * rng.uniform(64, *sb)
    L.config_set("splitTiming", 1)
    try:
        C, rep = zsplit_mult(A, B, "split_int8", 0, use3m, opA, opB)
    finally:
        L.config_set("splitTiming", 0)
    assert rep.precision == L.PREC["split_int8"]
    # one product launch covers every real product's moduli (3M: 3 products, 4M: 4)
    assert rep.launches == 1 and rep.products == (3 if use3m else 4) * rep.slices
    assert 0 <= rep.max_rel_err <= 1e-15 and rep.product_ms > 0
    assert zsampled_err(C, A, B, opA, opB) <= 1e-15


def test_complex_crt_gaussian_integers_exact():
    """Gaussian integers below 2^19 in each part: T1, T2, T3 are exact integer
    products and Re, Im are exact sums of them."""
    r = np.random.default_rng(7)
    A = (r.integers(-(1 << 19), 1 << 19, (70, 90)) + 1j * r.integers(-(1 << 19), 1 << 19, (70, 90)))
    B = (r.integers(-(1 << 19), 1 << 19, (90, 50)) + 1j * r.integers(-(1 << 19), 1 << 19, (90, 50)))
    C, rep = zsplit_mult(A, B, "split_int8", 0, True)
    Ar, Ai = A.real.astype(object), A.imag.astype(object)
    Br, Bi = B.real.astype(object), B.imag.astype(object)
    re = Ar.dot(Br) - Ai.dot(Bi)
    im = Ar.dot(Bi) + Ai.dot(Br)
    assert np.array_equal(C.real, np.array(re, dtype=np.float64))
    assert np.array_equal(C.imag, np.array(im, dtype=np.float64))
    assert rep.precision == L.PREC["split_int8"]


def same_nonfinite(C, ref):
    """C and ref have NaN, +Inf and -Inf in exactly the same places."""
    return (np.array_equal(np.isnan(C), np.isnan(ref)) and np.array_equal(np.isposinf(C), np.isposinf(ref))
            and np.array_equal(np.isneginf(C), np.isneginf(ref)))


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("prec", ["split_int8", "split_bf16", "split_fp16"])
def test_nonfinite_inputs_take_the_fp64_path(prec, cplx):
    """NaN and Inf entries (outside the rows the residual check samples) cannot be
    carried by residues or digit slices: the exponent pass flags them and the
    product is computed in FP64, whose NaN / +-Inf pattern is the reference
    arithmetic's; finite entries within the FP64 bound 2 K u (|A||B|)."""
    r = np.random.default_rng(17)
    M, N, K = 300, 257, 400
    A, B = r.standard_normal((M, K)), r.standard_normal((K, N))
    if cplx:
        A = A + 1j * r.standard_normal((M, K))
        B = B + 1j * r.standard_normal((K, N))
    A[123, 45] = np.inf
    A[201, 7] = np.nan
    B[300, 77] = -np.inf
    if cplx:
        C = f4j.CMatrix(A).mult(f4j.CMatrix(B), precision=prec).numpy()
        ref = o.cplx_gemm(A, B)
        parts = [(C.real, ref.real), (C.imag, ref.imag)]
        absA, absB = np.abs(A.real) + np.abs(A.imag), np.abs(B.real) + np.abs(B.imag)
    else:
        C, _ = split_mult(A, B, prec, 0)
        ref = o.real_reordered(A, B)
        parts = [(C, ref)]
        absA, absB = np.abs(A), np.abs(B)
    assert L.last_split_report().precision == L.PREC["fp64"]
    with np.errstate(invalid="ignore"):
        bound = (2 if cplx else 1) * 2 * K * 2.0**-53 * (np.where(np.isfinite(absA), absA, 0) @
                                                          np.where(np.isfinite(absB), absB, 0))
    for c, rf in parts:
        assert same_nonfinite(c, rf)
        fin = np.isfinite(rf)
        assert np.all(np.abs(c[fin] - rf[fin]) <= bound[fin])
