"""Host-side mirror of Flag4j's dense matrix types on the multiply path.

`Matrix`, `CMatrix` and `Shape` keep the reference's names, argument meaning
and error behaviour for the calls on the hot path, and run every product on
the GPU through the C ABI (include/flag4j_amd.h).  Citation prefix
M: = src/main/java/org/flag4j/.

Storage: `Matrix.data` is a flat row-major float64 array
(M:arrays/dense/Matrix.java:113-140); `CMatrix.data` a flat row-major
complex128 array, i.e. interleaved (re, im) doubles, which is how complex data
crosses the C ABI (the reference stores `Complex128[]`,
M:arrays/dense/CMatrix.java:96-110).
"""
import ctypes

import numpy as np

from . import _native

INT_MAX = 2**31 - 1


class LinearAlgebraException(Exception):
    """M:util/exceptions/LinearAlgebraException.java:31."""


class IllegalArgumentException(ValueError):
    """java.lang.IllegalArgumentException as thrown by M:util/ValidateParameters.java."""


class SingularMatrixException(ArithmeticError):
    """M:util/exceptions/SingularMatrixException.java: thrown by the back solver when the
    coefficient matrix is singular or near singular."""


class Shape:
    """Row-major shape (M:arrays/Shape.java:98-105, :188-193, :397-412)."""

    def __init__(self, *dims):
        if len(dims) == 1 and isinstance(dims[0], (tuple, list)):
            dims = tuple(dims[0])
        self.dims = tuple(int(d) for d in dims)
        total = 1
        for d in self.dims:
            if d < 0:
                raise IllegalArgumentException("All dimensions must be non-negative")
            total *= d
        if total > INT_MAX:       # Shape.java:405-407: Math.multiplyExact on int
            raise ArithmeticError("integer overflow")

    def get(self, i):
        return self.dims[i]

    def rank(self):
        return len(self.dims)

    def totalEntries(self):
        t = 1
        for d in self.dims:
            t *= d
        return t

    def __eq__(self, other):
        return isinstance(other, Shape) and self.dims == other.dims

    def __hash__(self):
        return hash(self.dims)

    def __repr__(self):
        return "(" + ", ".join(str(d) for d in self.dims) + ")"


def _mat_mult_shape_err(s1, s2):
    # M:util/ErrorMessages.java:52, :162-164
    return (f"Cannot multiply matrices with shapes {s1} and {s2}. Matrix multiplication requires the number "
            f"of columns in the first matrix to equal the number of rows in the second matrix.")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class _DenseMatrix:
    _dtype = np.float64

    def __init__(self, *args):
        if len(args) == 1:
            arr = np.array(args[0], dtype=self._dtype)
            if arr.ndim != 2:
                raise IllegalArgumentException("expected a 2D array of entries")
            self.shape = Shape(arr.shape)
            self._data = np.ascontiguousarray(arr).reshape(-1)
        elif len(args) == 2:
            shape, data = args
            shape = shape if isinstance(shape, Shape) else Shape(shape)
            data = np.ascontiguousarray(np.asarray(data, dtype=self._dtype).reshape(-1))
            if data.size != shape.totalEntries():   # AbstractDenseDoubleTensor.java:56-59
                raise IllegalArgumentException(
                    f"Shape {shape} does not match the number of entries {data.size}")
            self.shape, self._data = shape, data
        else:
            raise TypeError("expected (entries) or (shape, data)")

    # ---- accessors mirroring the reference's public fields ----
    @property
    def data(self):
        return self._data

    @property
    def numRows(self):
        return self.shape.get(0)

    @property
    def numCols(self):
        return self.shape.get(1)

    def dataLength(self):
        return self.data.size

    def numpy(self):
        return self.data.reshape(self.numRows, self.numCols)

    def T(self):
        """Materialised transpose (M:linalg/ops/TransposeDispatcher.java:77-99)."""
        return type(self)(Shape(self.numCols, self.numRows), self.numpy().T.copy())

    def copy(self):
        return type(self)(self.shape, self.data.copy())

    def __repr__(self):
        return f"{type(self).__name__}{self.shape}"


class Matrix(_DenseMatrix):
    """Real dense matrix (M:arrays/dense/Matrix.java)."""

    _dtype = np.float64

    @staticmethod
    def I(n):  # noqa: E743
        return Matrix(Shape(n, n), np.eye(n))

    def mult(self, b, precision="auto", slices=0):
        """C = this * b: Matrix.mult (M:arrays/dense/Matrix.java:638-640) via the GPU dispatcher;
        for a Vector b, Matrix.mult(Vector) (Matrix.java:1406-1408).

        Raises LinearAlgebraException if this.numCols != b.numRows
        (ReDeMatMultDispatcher.java:163-164; ValidateParameters.java:62-67)."""
        if isinstance(b, Vector):
            return self._mult_vector(b, precision)
        if isinstance(b, (CMatrix, CVector)):
            return _mixed_mult(self, b, precision, slices)
        if self.numCols != b.numRows:
            raise LinearAlgebraException(_mat_mult_shape_err(self.shape, b.shape))
        M, K, N = self.numRows, self.numCols, b.numCols
        out = np.empty(M * N, dtype=np.float64)
        _native.check(_native.load().f4j_dgemm(0, 0, M, N, K, _ptr(self.data), max(K, 1), _ptr(b.data),
                                               max(N, 1), _ptr(out), max(N, 1), _native.PREC[precision], slices))
        return Matrix(Shape(M, N), out)

    def _mult_vector(self, b, precision):
        """Matrix.mult(Vector) -> ReDeMatVecMultDispatcher.dispatch
        (M:linalg/ops/dispatch/ReDeMatVecMultDispatcher.java:141-181): LinearAlgebraException on a
        size mismatch (:147-148, ValidateParameters.ensureMatMultShapes)."""
        if self.numCols != b.size:
            raise LinearAlgebraException(_mat_mult_shape_err(self.shape, b.shape))
        M, K = self.numRows, self.numCols
        out = np.empty(M, dtype=np.float64)
        _native.check(_native.load().f4j_dgemv(M, K, _ptr(self.data), max(K, 1), _ptr(b.data), 1, _ptr(out), 1,
                                               _native.PREC[precision]))
        return Vector(out)

    def multTranspose(self, b, precision="auto", slices=0):
        """this * b^T without forming b^T (M:arrays/dense/Matrix.java:655-663).

        Raises IllegalArgumentException if this.numCols != b.numCols
        (ValidateParameters.ensureAllEqual, ValidateParameters.java:135-139)."""
        if self.numCols != b.numCols:
            raise IllegalArgumentException(
                f"Expecting values to be equal but got: {self.numCols}, {b.numCols}.")
        M, K, N = self.numRows, self.numCols, b.numRows
        out = np.empty(M * N, dtype=np.float64)
        _native.check(_native.load().f4j_dgemm(0, 1, M, N, K, _ptr(self.data), max(K, 1), _ptr(b.data),
                                               max(K, 1), _ptr(out), max(N, 1), _native.PREC[precision], slices))
        return Matrix(Shape(M, N), out)

    def pow(self, n):
        """Matrix.pow (M:arrays/dense/Matrix.java:1932-1950): exponentiation by squaring."""
        if self.numRows != self.numCols:
            raise LinearAlgebraException(_mat_mult_shape_err(self.shape, self.shape))
        if n < 0:
            raise IllegalArgumentException(f"Expecting value to be non-negative but got {n}.")
        if n == 0:
            return Matrix.I(self.numRows)
        if n == 1:
            return self.copy()
        if n == 2:
            return self.mult(self)
        result, base = Matrix.I(self.numRows), self
        while n > 0:
            if n & 1:
                result = result.mult(base)
            base = base.mult(base)
            n >>= 1
        return result

    def equals(self, other):
        """Matrix.equals (M:arrays/dense/Matrix.java:2102-2109 -> RealDenseEquals.tensorEquals):
        same shape and == elementwise, with NaN equal to NaN."""
        if not isinstance(other, Matrix) or self.shape != other.shape:
            return False
        a, b = self.data, other.data
        return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))

    __eq__ = equals
    __hash__ = None


class CMatrix(_DenseMatrix):
    """Complex128 dense matrix (M:arrays/dense/CMatrix.java).

    `H()` returns the conjugate transpose as a lazy operand: a following
    `mult` passes the original storage with op C to the GPU, so A^H*B never
    materialises A^H (the reference materialises it, AbstractDenseFieldMatrix.java:98-102)."""

    _dtype = np.complex128

    def __init__(self, *args, _parent=None):
        if _parent is not None:
            self._parent, self._op = _parent, "C"
            self.shape = Shape(_parent.numCols, _parent.numRows)
            self._data = None
            return
        self._parent, self._op = None, "N"
        super().__init__(*args)

    @property
    def data(self):
        if self._data is None:       # materialise a lazy H() on first access
            self._data = np.ascontiguousarray(self._parent.numpy().conj().T).reshape(-1)
        return self._data

    @staticmethod
    def I(n):  # noqa: E743
        return CMatrix(Shape(n, n), np.eye(n, dtype=np.complex128))

    def H(self):
        return CMatrix(_parent=self)

    def _operand(self):
        """(storage array, op code, stored row length) for the C ABI."""
        if self._op == "C" and self._data is None:
            p = self._parent
            return p.data, 2, max(p.numCols, 1)
        return self.data, 0, max(self.numCols, 1)

    def mult(self, b, precision="auto", slices=0, use_3m=False):
        """C = this * b: CMatrix.mult (M:arrays/dense/CMatrix.java:463-465) via the GPU dispatcher.

        Raises LinearAlgebraException on an inner-dimension mismatch
        (Cm128DeMatMultDispatcher.java:162-163).  Unlike the reference (whose
        kernels size dest as rows(a)*rows(b), Cm128DeMatMultKernels.java:47-77),
        any B shape is handled.  For a CVector b: CMatrix.mult(CVector)
        (M:arrays/backend/semiring_arrays/AbstractDenseSemiringMatrix.java:295-298)."""
        if isinstance(b, CVector):
            return self._mult_vector(b, precision)
        if isinstance(b, (Matrix, Vector)):
            return _mixed_mult(self, b, precision, slices)
        if self.numCols != b.numRows:
            raise LinearAlgebraException(_mat_mult_shape_err(self.shape, b.shape))
        M, K, N = self.numRows, self.numCols, b.numCols
        a_arr, opA, lda = self._operand()
        b_arr, opB, ldb = b._operand()
        out = np.empty(M * N, dtype=np.complex128)
        _native.check(_native.load().f4j_zgemm(opA, opB, M, N, K, _ptr(a_arr), lda, _ptr(b_arr), ldb,
                                               _ptr(out), max(N, 1), _native.PREC[precision], slices,
                                               int(use_3m)))
        return CMatrix(Shape(M, N), out)

    def _mult_vector(self, b, precision):
        """DenseSemiringMatMultDispatcher.dispatchVector (M:linalg/ops/dense/semiring_ops/
        DenseSemiringMatMultDispatcher.java:120-142): LinearAlgebraException on a size
        mismatch (ensureMatMultShapes, :124)."""
        if self.numCols != b.size:
            raise LinearAlgebraException(_mat_mult_shape_err(self.shape, Shape(b.size, 1)))
        M, K = self.numRows, self.numCols
        out = np.empty(M, dtype=np.complex128)
        _native.check(_native.load().f4j_zgemv(M, K, _ptr(self.data), max(K, 1), _ptr(b.data), 1, _ptr(out), 1,
                                               _native.PREC[precision]))
        return CVector(out)

    def multTranspose(self, b, precision="auto", slices=0):
        """this * b^T, transpose without conjugation
        (M:arrays/backend/semiring_arrays/AbstractDenseSemiringMatrix.java:330-334)."""
        if self.numCols != b.numCols:
            raise IllegalArgumentException(
                f"Expecting values to be equal but got: {self.numCols}, {b.numCols}.")
        M, K, N = self.numRows, self.numCols, b.numRows
        out = np.empty(M * N, dtype=np.complex128)
        _native.check(_native.load().f4j_zgemm(0, 1, M, N, K, _ptr(self.data), max(K, 1), _ptr(b.data),
                                               max(K, 1), _ptr(out), max(N, 1), _native.PREC[precision],
                                               slices, 0))
        return CMatrix(Shape(M, N), out)

    def pow(self, n):
        """CMatrix.pow (M:arrays/dense/CMatrix.java:844-866)."""
        if self.numRows != self.numCols:     # ValidateParameters.ensureSquare
            raise IllegalArgumentException(f"Expecting matrix to be square but got shape {self.shape}.")
        if n < 0:
            raise IllegalArgumentException(f"Expecting value to be non-negative but got {n}.")
        if n == 0:
            return CMatrix.I(self.numRows)
        if n == 1:
            return self.copy()
        if n == 2:
            return self.mult(self)
        result, base = CMatrix.I(self.numRows), self
        while n > 0:
            if n & 1:
                result = result.mult(base)
            base = base.mult(base)
            n >>= 1
        return result

    def equals(self, other):
        """CMatrix.equals (M:arrays/dense/CMatrix.java:1096-1103): Complex128.equals is
        re == re && im == im (M:numbers/Complex128.java:211-219)."""
        if not isinstance(other, CMatrix) or self.shape != other.shape:
            return False
        a, b = self.data, other.data
        return bool(np.all((a.real == b.real) & (a.imag == b.imag)))

    __eq__ = equals
    __hash__ = None


def _mixed_mult(a, b, precision, slices):
    """Real x complex products: Matrix.mult(CMatrix / CVector) (M:arrays/dense/Matrix.java:1793-1797,
    :1851-1853) and CMatrix.mult(Matrix / Vector) (M:arrays/dense/CMatrix.java:878-881, :894-896)
    -> MatrixMultiplyDispatcher (M:linalg/ops/MatrixMultiplyDispatcher.java:71-260), which raises
    LinearAlgebraException on an inner-dimension mismatch (ValidateParameters.ensureMatMultShapes)."""
    vec = isinstance(b, _DenseVector)
    K2, N = (b.size, 1) if vec else (b.numRows, b.numCols)
    if a.numCols != K2:
        raise LinearAlgebraException(_mat_mult_shape_err(a.shape, b.shape))
    M, K = a.numRows, a.numCols
    out = np.empty(M * N, dtype=np.complex128)
    f = _native.load().f4j_zdgemm if isinstance(a, CMatrix) else _native.load().f4j_dzgemm
    _native.check(f(M, N, K, _ptr(a.data), max(K, 1), _ptr(b.data), max(N, 1), _ptr(out), max(N, 1),
                    _native.PREC[precision], slices))
    return CVector(out) if vec else CMatrix(Shape(M, N), out)


class _DenseVector:
    _dtype = np.float64

    def __init__(self, entries):
        arr = np.ascontiguousarray(np.array(entries, dtype=self._dtype).reshape(-1))
        self.shape = Shape(arr.size)
        self._data = arr

    @property
    def data(self):
        return self._data

    @property
    def size(self):
        return self.shape.get(0)

    def numpy(self):
        return self._data

    def __repr__(self):
        return f"{type(self).__name__}{self.shape}"


class Vector(_DenseVector):
    """Real dense vector (M:arrays/dense/Vector.java)."""

    _dtype = np.float64

    def equals(self, other):
        """Vector.equals: same size and == elementwise, NaN equal to NaN (RealDenseEquals)."""
        if not isinstance(other, Vector) or self.shape != other.shape:
            return False
        a, b = self.data, other.data
        return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))

    __eq__ = equals
    __hash__ = None


class CVector(_DenseVector):
    """Complex128 dense vector (M:arrays/dense/CVector.java)."""

    _dtype = np.complex128

    def equals(self, other):
        """CVector.equals: Complex128.equals elementwise (M:numbers/Complex128.java:211-219)."""
        if not isinstance(other, CVector) or self.shape != other.shape:
            return False
        a, b = self.data, other.data
        return bool(np.all((a.real == b.real) & (a.imag == b.imag)))

    __eq__ = equals
    __hash__ = None


class RealLstsqSolver:
    """RealLstsqSolver (M:linalg/solvers/lstsq/RealLstsqSolver.java:131, LstsqSolver.java:157-181):
    X minimising ||A X - B|| through the reduced Householder QR of A and back substitution,
    run on the GPU in the reference's arithmetic (f4j_dlstsq).  ``precision`` applies to the
    product Q^H B only ("auto" as Matrix.mult; "strict" makes the whole solve bitwise equal
    to the reference's).

    Errors as the reference raises them: LinearAlgebraException when B's rows differ from A's
    (Qh.mult(B)), IllegalArgumentException when A has more columns than rows (R not square),
    SingularMatrixException when |det R| <= 2^-53 n or det R is NaN."""

    def __init__(self, precision="auto"):
        self.precision = precision

    def solve(self, A, b):
        if not isinstance(A, Matrix):
            raise TypeError("A must be a Matrix")
        M, N = A.numRows, A.numCols
        vector = isinstance(b, Vector)
        if not vector and not isinstance(b, Matrix):
            raise TypeError("b must be a Vector or a Matrix")
        rows = b.size if vector else b.numRows
        R = 1 if vector else b.numCols
        mn = min(M, N)
        if rows != M:                                      # Qh (mn x M) . b
            raise LinearAlgebraException(_mat_mult_shape_err(Shape(mn, M), Shape(rows, R)))
        if M < N or N == 0:                                # ensureSquare(R)
            raise IllegalArgumentException(f"Expecting matrix to be square but got shape {Shape(mn, N)}.")
        a = np.ascontiguousarray(A.data, dtype=np.float64).reshape(M, N)
        bb = np.ascontiguousarray(b.data, dtype=np.float64).reshape(M, R)
        x = np.empty((N, R))
        lib = _native.load()
        st = lib.f4j_dlstsq(M, N, R, _ptr(a), N, _ptr(bb), R, _ptr(x), R, _native.PREC[self.precision])
        if st == _native.ERR_SINGULAR:
            raise SingularMatrixException("Could not solve system.")
        _native.check(st)
        return Vector(x.reshape(-1)) if vector else Matrix(x)


class Invert:
    """Invert.pInv (M:linalg/Invert.java:292-298): the Moore-Penrose pseudo-inverse, on the GPU
    (f4j_dpinv).  The reference forms V S^-1 U^T from RealSVD; for the full-rank matrices it
    accepts (Invert.invDiag raises SingularMatrixException when |prod s_i| <= 2^-53 min(m, n))
    the pseudo-inverse is unique, and this computes it as R^-1 Q^T from the reduced Householder
    QR (a wide matrix through its transpose), in the reference's least-squares arithmetic:
    bit for bit RealLstsqSolver.solve(A, I).  Singular: |det R| = prod s_i, the same criterion."""

    @staticmethod
    def pInv(src):
        if not isinstance(src, Matrix):
            raise TypeError("pInv takes a Matrix")
        M, N = src.numRows, src.numCols
        if M == 0 or N == 0:       # the C ABI has no empty pseudo-inverse (F4J_ERR_SHAPE)
            raise IllegalArgumentException(f"pInv of an empty matrix of shape {src.shape}.")
        a = np.ascontiguousarray(src.data, dtype=np.float64).reshape(M, N)
        x = np.empty((N, M))
        st = _native.load().f4j_dpinv(M, N, _ptr(a), N, _ptr(x), M)
        if st == _native.ERR_SINGULAR:
            raise SingularMatrixException("Could not invert.")
        _native.check(st)
        return Matrix(x)
