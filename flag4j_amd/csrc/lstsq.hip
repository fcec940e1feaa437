// Least squares X = argmin ||A X - B|| in the reference's arithmetic
// (RealLstsqSolver.solve, M:linalg/solvers/lstsq/LstsqSolver.java:157-181):
//   1. reduced Householder QR of A, column by column
//      (M:linalg/decompositions/unitary/UnitaryDecomposition.java decompose,
//       M:linalg/decompositions/unitary/RealUnitaryDecomposition.java computeHouseholder /
//       updateData, M:linalg/transformations/Householder.java:156-188 leftMultReflector);
//   2. Q formed by applying the stored reflectors to I(M, N) last to first
//      (RealUnitaryDecomposition.java:147-168 getQ);
//   3. Y = Q^H B through the library's GEMM (the caller does that step);
//   4. back substitution R X = Y (M:linalg/solvers/exact/triangular/RealBackSolver.java solve).
//
// Every product and sum is rounded separately and every chain runs in the
// reference's order (this file is compiled without floating-point contraction),
// so T, Q, R and X equal the reference's bit for bit.  Parallelism comes from the
// columns: the reflector's inner products v^T S are one k-ascending chain per
// column of S (one lane each), the rank-one update is element-parallel, and the
// back substitution runs one lane per right-hand side.  The chains that remain
// serial (the reflector norm, each column's v^T S, each back-substitution row) are
// the reference's own dependence chains.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "f4j_internal.h"

namespace f4j {

namespace {

constexpr int HV_THREADS = 1024;        // reflector of one column: one workgroup
constexpr int DOT_THREADS = 64;         // one wave per 64 columns of S
constexpr int UPD_COLS = 64, UPD_ROWS = 4, UPD_ROWS_PER_THREAD = 8;

// java.lang.Math.max on non-negative operands: NaN if either is NaN.
__device__ inline double nan_max(double a, double b) {
    if (a != a || b != b) return __longlong_as_double(0x7ff8000000000000ll);
    return a > b ? a : b;
}

// Reflector of column j of T (M x ldt, rows j..M-1): v (absolute row index, v[j] = 1),
// qf[j] (0 when the column is below EPS_F64 and no reflector applies) and
// diag[j] = -norm, the value R[j][j] takes.
__global__ __launch_bounds__(HV_THREADS) void hh_vector_kernel(const double *__restrict__ T, int64_t ldt, int64_t M,
                                                               int64_t j, double *__restrict__ v,
                                                               double *__restrict__ qf, double *__restrict__ diag) {
    __shared__ double red[HV_THREADS / 64];
    __shared__ double inv_shift;
    const int tid = threadIdx.x;
    double m = 0.0;
    for (int64_t i = j + tid; i < M; i += HV_THREADS) m = nan_max(fabs(T[i * ldt + j]), m);
    for (int off = 32; off > 0; off >>= 1) m = nan_max(__shfl_xor(m, off, 64), m);
    if ((tid & 63) == 0) red[tid >> 6] = m;
    __syncthreads();
    double max_abs = red[0];
    for (int w = 1; w < HV_THREADS / 64; ++w) max_abs = nan_max(red[w], max_abs);
    if (!(max_abs >= 0x1.0p-53)) {                       // no reflector (also when max_abs is NaN)
        if (tid == 0) {
            qf[j] = 0.0;
            diag[j] = 0.0;
        }
        return;
    }
    const double max_inv = 1.0 / max_abs;
    for (int64_t i = j + tid; i < M; i += HV_THREADS) v[i] = T[i * ldt + j] * max_inv;
    __syncthreads();
    if (tid == 0) {
        double norm = 0.0;                               // one k-ascending chain, as computePhasedNorm
        for (int64_t i = j; i < M; ++i) {
            const double s = v[i];
            norm = norm + s * s;
        }
        norm = sqrt(norm);
        const double vj = v[j];
        if (vj < 0) norm = -norm;
        const double shift = vj + norm;
        qf[j] = shift / norm;
        norm = norm * max_abs;
        diag[j] = -norm;
        inv_shift = 1.0 / shift;
        v[j] = 1.0;
    }
    __syncthreads();
    const double is = inv_shift;
    for (int64_t i = j + 1 + tid; i < M; i += HV_THREADS) v[i] = v[i] * is;
}

// w[c] = alpha (S[r0][c] + sum_{k = r0+1}^{r1-1} v[k] S[k][c]) for c in [c0, ncols): the
// reflector's leading entry is 1, the sum is one k-ascending chain per column, and
// alpha = qf[j] scales afterwards (Householder.leftMultReflector).  Nothing runs when
// qf[j] is 0.
__global__ __launch_bounds__(DOT_THREADS) void hh_dot_kernel(const double *__restrict__ S, int64_t lds, int64_t c0,
                                                             int64_t ncols, int64_t r0, int64_t r1,
                                                             const double *__restrict__ v, int64_t incv,
                                                             const double *__restrict__ qf_j,
                                                             double *__restrict__ w) {
    const double alpha = *qf_j;
    if (alpha == 0.0) return;
    const int64_t c = c0 + (int64_t)blockIdx.x * DOT_THREADS + threadIdx.x;
    if (c >= ncols) return;
    const double *col = S + c;
    double acc = 1.0 * col[r0 * lds];
    int64_t k = r0 + 1;
    for (; k + 4 <= r1; k += 4) {
        const double p0 = v[k * incv] * col[k * lds];
        const double p1 = v[(k + 1) * incv] * col[(k + 1) * lds];
        const double p2 = v[(k + 2) * incv] * col[(k + 2) * lds];
        const double p3 = v[(k + 3) * incv] * col[(k + 3) * lds];
        acc = acc + p0;
        acc = acc + p1;
        acc = acc + p2;
        acc = acc + p3;
    }
    for (; k < r1; ++k) acc = acc + v[k * incv] * col[k * lds];
    w[c] = acc * alpha;
}

// S[i][c] -= v[i] w[c] for i in [r0, r1), c in [c0, ncols) (v[r0] taken as 1).  With
// diag (the factorisation), column c0 then holds the stored reflector instead:
// diag[j] at row r0 and v below it (RealUnitaryDecomposition.updateData).
__global__ __launch_bounds__(UPD_COLS *UPD_ROWS) void hh_update_kernel(double *__restrict__ S, int64_t lds,
                                                                       int64_t c0, int64_t ncols, int64_t r0,
                                                                       int64_t r1, const double *__restrict__ v,
                                                                       int64_t incv,
                                                                       const double *__restrict__ qf_j,
                                                                       const double *__restrict__ w,
                                                                       const double *__restrict__ diag_j) {
    if (*qf_j == 0.0) return;
    const int64_t c = c0 + (int64_t)blockIdx.x * UPD_COLS + threadIdx.x;
    if (c >= ncols) return;
    const double wc = w[c];
    const bool stored = diag_j != nullptr && c == c0;
    const int64_t ib = r0 + ((int64_t)blockIdx.y * UPD_ROWS + threadIdx.y) * UPD_ROWS_PER_THREAD;
#pragma unroll
    for (int t = 0; t < UPD_ROWS_PER_THREAD; ++t) {
        const int64_t i = ib + t;
        if (i >= r1) break;
        const double vi = i == r0 ? 1.0 : v[i * incv];
        double *s = S + i * lds + c;
        if (stored)
            *s = i == r0 ? *diag_j : vi;
        else
            *s = *s - vi * wc;
    }
}

// Q = I(M, N) before the reflectors are applied (Matrix.I(numRows, qCols)).
__global__ void eye_kernel(double *__restrict__ Q, int64_t M, int64_t N, int64_t ldq) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * N) return;
    const int64_t i = idx / N, c = idx % N;
    Q[i * ldq + c] = i == c ? 1.0 : 0.0;
}

// Qh (N x M) = Q^T (Q is M x N): the materialised Q.H() of a real Q.
__global__ __launch_bounds__(256) void transpose_kernel(const double *__restrict__ Q, int64_t M, int64_t N,
                                                         int64_t ldq, double *__restrict__ Qh, int64_t ldh) {
    __shared__ double tile[32][33];
    const int64_t i0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;     // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = i0 + r, c = c0 + tx;
        if (i < M && c < N) tile[r][tx] = Q[i * ldq + c];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t c = c0 + r, i = i0 + tx;
        if (c < N && i < M) Qh[c * ldh + i] = tile[tx][r];
    }
}

// RealBackSolver.solve(U, B): X[n-1] = Y[n-1] / U[n-1][n-1], then for i = n-2 .. 0
// X[i] = (Y[i] - sum_{k > i} U[i][k] X[k]) / U[i][i] with the sum one k-ascending
// chain from 0; one lane per right-hand side.  U is the upper part of T (ld ldt).
__global__ __launch_bounds__(64) void back_solve_kernel(const double *__restrict__ T, int64_t ldt, int64_t n,
                                                        const double *__restrict__ Y, int64_t ldy, double *X,
                                                        int64_t ldx, int64_t R) {
    const int64_t c = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (c >= R) return;
    X[(n - 1) * ldx + c] = Y[(n - 1) * ldy + c] / T[(n - 1) * ldt + (n - 1)];
    for (int64_t i = n - 2; i >= 0; --i) {
        const double *u = T + i * ldt;
        double sum = 0.0;
        int64_t k = i + 1;
        for (; k + 4 <= n; k += 4) {
            const double p0 = u[k] * X[k * ldx + c];
            const double p1 = u[k + 1] * X[(k + 1) * ldx + c];
            const double p2 = u[k + 2] * X[(k + 2) * ldx + c];
            const double p3 = u[k + 3] * X[(k + 3) * ldx + c];
            sum = sum + p0;
            sum = sum + p1;
            sum = sum + p2;
            sum = sum + p3;
        }
        for (; k < n; ++k) sum = sum + u[k] * X[k * ldx + c];
        X[i * ldx + c] = (Y[i * ldy + c] - sum) / u[i];
    }
}

// det R as the back solver forms it: U[n-1][n-1], then times U[i][i] for
// i = n-2 .. 0 when there is at least one right-hand side.
__global__ void det_kernel(const double *__restrict__ T, int64_t ldt, int64_t n, int64_t R, double *__restrict__ det) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double d = T[(n - 1) * ldt + (n - 1)];
    if (R > 0)
        for (int64_t i = n - 2; i >= 0; --i) d = d * T[i * ldt + i];
    *det = d;
}

dim3 update_grid(int64_t cols, int64_t rows) {
    const int64_t rows_per_block = (int64_t)UPD_ROWS * UPD_ROWS_PER_THREAD;
    return dim3((unsigned)((cols + UPD_COLS - 1) / UPD_COLS), (unsigned)((rows + rows_per_block - 1) / rows_per_block));
}

}  // namespace

hipError_t launch_householder_qr(double *T, int64_t M, int64_t N, int64_t ldt, double *v, double *w, double *qf,
                                 double *diag, hipStream_t stream) {
    const int64_t mn = M < N ? M : N;
    for (int64_t j = 0; j < mn; ++j) {
        hh_vector_kernel<<<1, HV_THREADS, 0, stream>>>(T, ldt, M, j, v, qf, diag);
        const int64_t cols = N - j;
        hh_dot_kernel<<<(unsigned)((cols + DOT_THREADS - 1) / DOT_THREADS), DOT_THREADS, 0, stream>>>(
            T, ldt, j, N, j, M, v, 1, qf + j, w);
        hh_update_kernel<<<update_grid(cols, M - j), dim3(UPD_COLS, UPD_ROWS), 0, stream>>>(T, ldt, j, N, j, M, v, 1,
                                                                                            qf + j, w, diag + j);
    }
    return hipGetLastError();
}

hipError_t launch_householder_q(const double *T, int64_t M, int64_t N, int64_t ldt, const double *qf, double *Q,
                                int64_t ldq, double *w, hipStream_t stream) {
    const int64_t mn = M < N ? M : N;
    if (M * mn > 0) eye_kernel<<<(unsigned)((M * mn + 255) / 256), 256, 0, stream>>>(Q, M, mn, ldq);
    for (int64_t j = mn - 1; j >= 0; --j) {
        const int64_t cols = mn - j;
        hh_dot_kernel<<<(unsigned)((cols + DOT_THREADS - 1) / DOT_THREADS), DOT_THREADS, 0, stream>>>(
            Q, ldq, j, mn, j, M, T + j, ldt, qf + j, w);
        hh_update_kernel<<<update_grid(cols, M - j), dim3(UPD_COLS, UPD_ROWS), 0, stream>>>(
            Q, ldq, j, mn, j, M, T + j, ldt, qf + j, w, nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_transpose(const double *Q, int64_t M, int64_t N, int64_t ldq, double *Qh, int64_t ldh,
                            hipStream_t stream) {
    if (M == 0 || N == 0) return hipSuccess;
    transpose_kernel<<<dim3((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32)), 256, 0, stream>>>(Q, M, N, ldq, Qh,
                                                                                                    ldh);
    return hipGetLastError();
}

hipError_t launch_back_solve(const double *T, int64_t ldt, int64_t n, const double *Y, int64_t ldy, double *X,
                             int64_t ldx, int64_t R, double *det, hipStream_t stream) {
    if (R > 0) back_solve_kernel<<<(unsigned)((R + 63) / 64), 64, 0, stream>>>(T, ldt, n, Y, ldy, X, ldx, R);
    det_kernel<<<1, 64, 0, stream>>>(T, ldt, n, R, det);
    return hipGetLastError();
}

}  // namespace f4j
