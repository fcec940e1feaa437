"""GPU parity of the least-squares solve (f4j_dlstsq / f4j_dlstsq_dev, RealLstsqSolver.solve)
against the oracle's restatement (Householder QR, Q, Q^H B, back substitution) and the
reference's golden vectors (T:linalg/solvers/RealLstsqSolverTests.java:17-98), and of the
pseudo-inverse built on it (f4j_dpinv, Invert.pInv).

Bars:
  * STRICT precision: X and the status equal the oracle's bit for bit (the QR, Q and the
    back substitution always run in the reference's order; STRICT makes the product too);
  * AUTO where the reference's product order is what AUTO runs (one right-hand side, the
    small path): bit for bit as well;
  * FP64 (MFMA) product: max |X - X_ref| / max |X_ref| <= 1e-10 on well-conditioned systems
    (the product's error, within 2 K u |Q||B|, carried through the back substitution).
"""
import ctypes

import numpy as np
import pytest

import flag4j_amd as f4j
from flag4j_amd import _native, device
from conftest import real_mat
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def p(a):
    return ctypes.c_void_p(a.ctypes.data)


def gpu_lstsq(A, B, precision="auto"):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B2 = np.ascontiguousarray(B.reshape(B.shape[0], -1), dtype=np.float64)
    M, N = A.shape
    R = B2.shape[1]
    X = np.full((N, max(R, 1)), np.nan)
    st = _native.load().f4j_dlstsq(M, N, R, p(A), N, p(B2), max(R, 1), p(X), max(R, 1), _native.PREC[precision])
    return st, X[:, :R]


def same_bits(X, Y):
    return np.array_equal(np.ascontiguousarray(X).view(np.int64), np.ascontiguousarray(Y).view(np.int64))


def system(rng, M, N, R, spread=2):
    """A with columns scaled over 10^-spread .. 10^spread (a Householder pass sees every
    magnitude) and B with a component outside A's range."""
    A = rng.standard_normal((M, N)) * 10.0 ** rng.uniform(-spread, spread, N)
    B = rng.standard_normal((M, R))
    return A, B


@pytest.mark.parametrize("case", ["lstsq_vec_1", "lstsq_vec_2", "lstsq_mat_1", "lstsq_mat_2"])
def test_golden_vectors_through_the_solver(golden, case):
    """RealLstsqSolverTests.java:17-98: solve(A, b) -> Vector, solve(A, B) -> Matrix, compared
    with assertEquals (Arrays.equals of the entries: bitwise)."""
    g = golden[case]
    A, B, exp = real_mat(g["a"]), real_mat(g["b"]), real_mat(g["x"])
    solver = f4j.RealLstsqSolver()
    if case.startswith("lstsq_vec"):
        x = solver.solve(f4j.Matrix(A), f4j.Vector(B.reshape(-1)))
        assert isinstance(x, f4j.Vector)
    else:
        x = solver.solve(f4j.Matrix(A), f4j.Matrix(B))
        assert isinstance(x, f4j.Matrix)
    assert same_bits(np.asarray(x.data).reshape(-1), exp.reshape(-1))


@pytest.mark.parametrize("M,N,R", [(50, 50, 1), (60, 40, 3), (300, 120, 7), (1000, 64, 1), (257, 255, 3),
                                   (129, 1, 2), (700, 300, 40), (1500, 700, 1), (64, 64, 64), (3000, 17, 5)])
def test_strict_bitwise(M, N, R):
    rng = np.random.default_rng(M * 7 + N * 3 + R)
    A, B = system(rng, M, N, R)
    st, X = gpu_lstsq(A, B, "strict")
    ost, OX = o.lstsq(A, B)
    assert st == _native.OK and ost == 0
    assert same_bits(X, OX), np.max(np.abs(X - OX))


@pytest.mark.parametrize("M,N", [(40, 30), (2000, 300), (999, 998)])
def test_auto_one_rhs_is_bitwise(M, N):
    """One right-hand side: Qh.mult(b) is a matrix-vector product, which AUTO runs in the
    reference's order (the k-ascending chain below the small threshold, 64-wide blocks above)."""
    rng = np.random.default_rng(M + N)
    A, b = system(rng, M, N, 1)
    st, X = gpu_lstsq(A, b, "auto")
    assert st == _native.OK
    assert same_bits(X, o.lstsq(A, b)[1])


def test_auto_small_path_is_bitwise():
    rng = np.random.default_rng(5)
    A, B = system(rng, 30, 20, 4)                 # Qh (20 x 30), B (30 x 4): 720 entries < 2450
    st, X = gpu_lstsq(A, B, "auto")
    assert st == _native.OK and same_bits(X, o.lstsq(A, B)[1])


def test_strict_bitwise_at_size():
    rng = np.random.default_rng(11)
    A, B = system(rng, 2048, 1024, 8, spread=1)
    st, X = gpu_lstsq(A, B, "strict")
    ost, OX = o.lstsq(A, B)
    assert st == _native.OK and ost == 0 and same_bits(X, OX)


@pytest.mark.parametrize("precision", ["auto", "fp64", "split_int8"])
def test_mfma_product_close(precision):
    rng = np.random.default_rng(17)
    A, B = system(rng, 1536, 512, 48, spread=0)
    st, X = gpu_lstsq(A, B, precision)
    ost, OX = o.lstsq(A, B)
    assert st == _native.OK and ost == 0
    assert np.max(np.abs(X - OX)) <= 1e-10 * np.max(np.abs(OX))


def test_singular_and_rank_deficient():
    """A zero column leaves R[j][j] = 0, so det R = 0: SingularMatrixException.  A duplicated
    column leaves R[j][j] at rounding level, far above 2^-53 n in |det R|: the reference
    solves it, and so does the library, with the same bits."""
    rng = np.random.default_rng(23)
    A, B = system(rng, 200, 60, 2, spread=0)
    Z = A.copy()
    Z[:, 17] = 0.0
    st, _ = gpu_lstsq(Z, B, "strict")
    assert st == _native.ERR_SINGULAR and o.lstsq(Z, B)[0] == 1
    assert b"Could not solve system" in _native.load().f4j_last_error()
    with pytest.raises(f4j.SingularMatrixException):
        f4j.RealLstsqSolver().solve(f4j.Matrix(Z), f4j.Matrix(B))
    D = A.copy()
    D[:, 17] = D[:, 16]
    st, X = gpu_lstsq(D, B, "strict")
    ost, OX = o.lstsq(D, B)
    assert st == _native.OK and ost == 0 and same_bits(X, OX)


def test_nan_and_tiny_columns():
    """A NaN entry makes its column's max NaN (no reflector) and det R NaN: singular, as the
    oracle says.  A column below 2^-53 gets no reflector either (qf = 0) and is carried as it is."""
    rng = np.random.default_rng(29)
    A, B = system(rng, 90, 30, 2, spread=0)
    An = A.copy()
    An[40, 7] = np.nan
    assert gpu_lstsq(An, B, "strict")[0] == _native.ERR_SINGULAR and o.lstsq(An, B)[0] == 1
    At = A.copy()
    At[:, 29] = 0.0
    At[29:, 29] = 1e-17                           # after 29 reflectors the tail is still tiny
    st, X = gpu_lstsq(At, B, "strict")
    ost, OX = o.lstsq(At, B)
    assert st == (_native.OK if ost == 0 else _native.ERR_SINGULAR)
    assert same_bits(X, OX)


def test_no_right_hand_side():
    """R = 0: nothing to solve; the back solver still checks det R (its last diagonal entry)."""
    rng = np.random.default_rng(31)
    A, _ = system(rng, 20, 10, 1)
    st, X = gpu_lstsq(A, np.zeros((20, 0)))
    assert st == _native.OK and X.shape == (10, 0)


def test_device_entry_with_leading_dimensions():
    rng = np.random.default_rng(37)
    M, N, R, lda, ldb, ldx = 400, 150, 6, 153, 9, 11
    A, B = system(rng, M, N, R)
    Ap = np.zeros((M, lda))
    Ap[:, :N] = A
    Bp = np.zeros((M, ldb))
    Bp[:, :R] = B
    L = _native.load()
    with device.DeviceBuffer.from_array(Ap) as dA, device.DeviceBuffer.from_array(Bp) as dB, \
            device.DeviceBuffer(N * ldx * 8) as dX:
        dX.upload(np.full(N * ldx, 7.0))
        _native.check(L.f4j_dlstsq_dev(M, N, R, dA.ptr, lda, dB.ptr, ldb, dX.ptr, ldx, _native.PREC["strict"], None))
        X = dX.download(np.float64, N * ldx).reshape(N, ldx)
    assert same_bits(np.ascontiguousarray(X[:, :R]), o.lstsq(A, B)[1])
    assert np.all(X[:, R:] == 7.0)


# ---- pseudo-inverse (f4j_dpinv, Invert.pInv) ----

def gpu_pinv(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    M, N = A.shape
    X = np.full((N, M), np.nan)
    st = _native.load().f4j_dpinv(M, N, p(A), N, p(X), M)
    return st, X


def penrose_residuals(A, X):
    """The four Moore-Penrose conditions, each relative to the size of its terms."""
    nA, nX = np.linalg.norm(A, 2), np.linalg.norm(X, 2)
    AX, XA = A @ X, X @ A
    return (np.linalg.norm(AX @ A - A, 2) / nA, np.linalg.norm(XA @ X - X, 2) / nX,
            np.linalg.norm(AX - AX.T, 2), np.linalg.norm(XA - XA.T, 2))


@pytest.mark.parametrize("M,N", [(1, 1), (7, 7), (130, 40), (40, 130), (300, 299), (257, 512), (1000, 64)])
def test_pinv_equals_the_reference_least_squares_arithmetic(M, N):
    """Bit for bit RealLstsqSolver.solve(A, I) of the oracle (its QR, Q, Q^H I and back
    substitution; Q^H I = Q^H exactly) for a tall A, and its transpose through A^T for a wide A."""
    rng = np.random.default_rng(100 + M + N)
    A, _ = system(rng, M, N, 1)
    st, X = gpu_pinv(A)
    if M >= N:
        ost, OX = o.lstsq(A, np.eye(M))
    else:
        ost, OXt = o.lstsq(np.ascontiguousarray(A.T), np.eye(N))
        OX = np.ascontiguousarray(OXt.T)
    assert st == _native.OK and ost == 0
    assert same_bits(X, OX)


@pytest.mark.parametrize("M,N", [(64, 64), (500, 120), (120, 500), (1024, 512)])
def test_pinv_moore_penrose_conditions(M, N):
    """What the mathematics fixes: A^+ is the unique X with A X A = A, X A X = X and A X, X A
    symmetric.  On matrices with condition number 1e2 every residual is a small multiple of
    the rounding unit; numpy's SVD-based pinv (LAPACK), an independent computation of the same
    unique matrix, agrees within the conditioning's bound."""
    rng = np.random.default_rng(7 * M + N)
    U, _ = np.linalg.qr(rng.standard_normal((M, min(M, N))))
    V, _ = np.linalg.qr(rng.standard_normal((N, min(M, N))))
    # condition number 100, prod s_i = 1: the reference's singularity test is on |det R| =
    # prod s_i <= 2^-53 min(M, N), which values spread below 1 alone would trip
    s = np.logspace(1, -1, min(M, N))
    A = (U * s) @ V.T
    st, X = gpu_pinv(A)
    assert st == _native.OK
    u = 2.0**-53
    bound = 64 * max(M, N) * u * 100.0
    r = penrose_residuals(A, X)
    assert max(r[0], r[2], r[3]) <= bound and r[1] <= bound * 100.0, r
    ref = np.linalg.pinv(A)
    assert np.linalg.norm(X - ref, 2) <= bound * 100.0 * np.linalg.norm(ref, 2)
    Y = f4j.Invert.pInv(f4j.Matrix(A)).numpy()
    assert same_bits(Y, X)


def test_pinv_singular_matrices():
    """A zero column (tall) or a zero row (wide) makes det R = 0: SingularMatrixException, as
    Invert.invDiag raises it for a zero singular value (|prod s_i| = |det R|)."""
    rng = np.random.default_rng(41)
    A, _ = system(rng, 90, 30, 1, spread=0)
    Z = A.copy()
    Z[:, 11] = 0.0
    assert gpu_pinv(Z)[0] == _native.ERR_SINGULAR
    assert gpu_pinv(np.ascontiguousarray(Z.T))[0] == _native.ERR_SINGULAR
    assert b"Could not invert" in _native.load().f4j_last_error()
    with pytest.raises(f4j.SingularMatrixException):
        f4j.Invert.pInv(f4j.Matrix(Z))


def test_pinv_device_entry_with_leading_dimensions():
    rng = np.random.default_rng(43)
    for M, N in ((300, 70), (70, 300)):
        lda, ldx = N + 3, M + 5
        A, _ = system(rng, M, N, 1)
        Ap = np.zeros((M, lda))
        Ap[:, :N] = A
        L = _native.load()
        with device.DeviceBuffer.from_array(Ap) as dA, device.DeviceBuffer(N * ldx * 8) as dX:
            dX.upload(np.full(N * ldx, 7.0))
            _native.check(L.f4j_dpinv_dev(M, N, dA.ptr, lda, dX.ptr, ldx, None))
            X = dX.download(np.float64, N * ldx).reshape(N, ldx)
        assert same_bits(np.ascontiguousarray(X[:, :M]), gpu_pinv(A)[1])
        assert np.all(X[:, M:] == 7.0)
