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
// so T, Q, R and X equal the reference's bit for bit.  The serial parts are the
// reference's own dependence chains -- the reflector norm, each column's v^T S,
// each back-substitution row -- and each runs on one wave (wave_chain): the lanes
// fetch and multiply 64 terms at a time with loads several chunks ahead, the
// additions follow in order.  Parallelism comes from the columns (a wave per
// column of v^T S, a wave per right-hand side) and from the element-parallel
// rank-one updates.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "f4j_internal.h"

namespace f4j {

namespace {

constexpr int HV_THREADS = 1024;        // reflector of one column: one workgroup
constexpr int CW = 4;                   // columns of S per workgroup of the inner-product kernel (a wave
                                        // each, one per SIMD: the additions of a chain keep a SIMD busy)
constexpr int BW = 4;                   // right-hand sides per workgroup of the back substitution (a wave each)
constexpr int UPD_COLS = 64, UPD_ROWS = 4, UPD_ROWS_PER_THREAD = 8;
constexpr int CHAIN_AHEAD = 4;          // 64-term chunks of a chain in flight ahead of the one being summed

// java.lang.Math.max on non-negative operands: NaN if either is NaN.
__device__ inline double nan_max(double a, double b) {
    if (a != a || b != b) return __longlong_as_double(0x7ff8000000000000ll);
    return a > b ? a : b;
}

// One wave sums n products in order: acc + t(0) + t(1) + ... + t(n-1), each
// t(e) = a(e) * b(e) rounded, each addition rounded, left to right -- the
// reference's dependence chain.  load(e, a, b) fetches the two factors of term e.
//
// The lanes fetch and multiply 64 terms (a chunk) at a time; CHAIN_AHEAD chunks
// of loads stay in flight (a ring of registers; fetches past the end read term
// n-1 again and are never added).  Chunk c is handed to the chain through half
// c & 1 of `buf` (128 doubles of this wave's LDS, 16-byte aligned), written one
// chunk ahead, and read back in groups of 8 terms with 3 groups in flight ahead
// of the additions, across chunk boundaries, so neither the loads nor the LDS
// round trip stall the chain.  The additions run in every lane; all lanes end
// with the same acc.
template <class Load>
__device__ inline double wave_chain(double acc, int64_t n, Load load, double *buf) {
    if (n <= 0) return acc;
    const int lane = threadIdx.x & 63;
    const int64_t nch = (n + 63) >> 6, nfull = n >> 6, last = n - 1;
    auto fetch = [&](int64_t chunk, double &a, double &b) {
        const int64_t e = chunk * 64 + lane;
        load(e < last ? e : last, a, b);
    };
    static_assert(CHAIN_AHEAD % 2 == 0, "chunk c uses ring slot c % CHAIN_AHEAD and LDS half c % 2");
    double ra[CHAIN_AHEAD], rb[CHAIN_AHEAD];
#pragma unroll
    for (int d = 0; d < CHAIN_AHEAD; ++d) fetch(d, ra[d], rb[d]);
    // chunk 0 into half 0; the ring then holds chunks 1 .. CHAIN_AHEAD
    buf[lane] = ra[0] * rb[0];
    fetch(CHAIN_AHEAD, ra[0], rb[0]);
    __builtin_amdgcn_wave_barrier();
    double2 g[4][4];                                   // group window: group t of a chunk in g[t % 4]
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) g[t][i] = reinterpret_cast<const double2 *>(buf)[t * 4 + i];

    // Full chunk c (half h): stage chunk c + 1 (ring slot sn) into the other half,
    // then add chunk c's 64 terms while reading ahead into chunk c + 1.
    auto full_chunk = [&](const int h, double &a_next, double &b_next, int64_t refill) {
        const double2 *cur = reinterpret_cast<const double2 *>(buf + h * 64);
        const double2 *nxt = reinterpret_cast<const double2 *>(buf + (h ^ 1) * 64);
        buf[(h ^ 1) * 64 + lane] = a_next * b_next;
        if (refill >= 0) fetch(refill, a_next, b_next);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                acc = acc + g[t % 4][i].x;
                acc = acc + g[t % 4][i].y;
            }
            __builtin_amdgcn_sched_barrier(0);
            const int u = t + 3;
#pragma unroll
            for (int i = 0; i < 4; ++i) g[u % 4][i] = u < 8 ? cur[u * 4 + i] : nxt[(u - 8) * 4 + i];
        }
        __builtin_amdgcn_wave_barrier();
    };

    int64_t q = 0;
    for (; q + CHAIN_AHEAD <= nfull; q += CHAIN_AHEAD) {
#pragma unroll
        for (int d = 0; d < CHAIN_AHEAD; ++d) {
            const int sn = (d + 1) % CHAIN_AHEAD;
            full_chunk(d & 1, ra[sn], rb[sn], q + d + 1 + CHAIN_AHEAD);
        }
    }
    // the last chunks: fewer than CHAIN_AHEAD full ones, then possibly a partial one
    // (chunk q + d sits in ring slot d and half d & 1; the ring needs no refill)
#pragma unroll
    for (int d = 0; d < CHAIN_AHEAD; ++d) {
        if (q + d >= nch) break;
        const int sn = (d + 1) % CHAIN_AHEAD;
        if (q + d < nfull) {
            full_chunk(d & 1, ra[sn], rb[sn], -1);
        } else {
            const double *cur = buf + (d & 1) * 64;
            const int left = (int)(n - (q + d) * 64);
            for (int l = 0; l < left; ++l) acc = acc + cur[l];
        }
    }
    return acc;
}

// Reflector of column j of T (M x ldt, rows j..M-1): v (absolute row index, v[j] = 1),
// qf[j] (0 when the column is below EPS_F64 and no reflector applies) and
// diag[j] = -norm, the value R[j][j] takes.
__global__ __launch_bounds__(HV_THREADS) void hh_vector_kernel(const double *__restrict__ T, int64_t ldt, int64_t M,
                                                               int64_t j, double *v, double *__restrict__ qf,
                                                               double *__restrict__ diag) {
    __shared__ double red[HV_THREADS / 64];
    __shared__ __align__(16) double stage[128];
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
    if (tid < 64) {
        // the norm as computePhasedNorm sums it: one chain of squares, i ascending from 0
        const double *vj = v + j;
        double norm = wave_chain(0.0, M - j, [&](int64_t e, double &a, double &b) { a = b = vj[e]; }, stage);
        if (tid == 0) {
            norm = sqrt(norm);
            const double v0 = v[j];
            if (v0 < 0) norm = -norm;
            const double shift = v0 + norm;
            qf[j] = shift / norm;
            norm = norm * max_abs;
            diag[j] = -norm;
            inv_shift = 1.0 / shift;
        }
    }
    __syncthreads();
    const double is = inv_shift;
    for (int64_t i = j + 1 + tid; i < M; i += HV_THREADS) v[i] = v[i] * is;
    if (tid == 0) v[j] = 1.0;
}

// The stored reflector of column j as a contiguous vector for Q's formation:
// v[j] = 1, v[i] = T[i][j] below (RealUnitaryDecomposition.getQ).
__global__ void hh_gather_kernel(const double *__restrict__ T, int64_t ldt, int64_t M, int64_t j,
                                 double *__restrict__ v) {
    const int64_t i = j + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) v[i] = i == j ? 1.0 : T[i * ldt + j];
}

// w[c] = alpha (S[r0][c] + sum_{k = r0+1}^{r1-1} v[k] S[k][c]) for c in [c0, ncols): the
// reflector's leading entry is 1, the sum is one k-ascending chain per column (a
// wave each), and alpha = qf[j] scales afterwards (Householder.leftMultReflector).
// Nothing runs when qf[j] is 0.
__global__ __launch_bounds__(64 * CW) void hh_dot_kernel(const double *__restrict__ S, int64_t lds, int64_t c0,
                                                         int64_t ncols, int64_t r0, int64_t r1,
                                                         const double *__restrict__ v,
                                                         const double *__restrict__ qf_j, double *__restrict__ w) {
    __shared__ __align__(16) double stage[CW][128];
    const double alpha = *qf_j;
    if (alpha == 0.0) return;
    const int wave = threadIdx.x >> 6;
    const int64_t c = c0 + (int64_t)blockIdx.x * CW + wave;
    if (c >= ncols) return;
    const double *col = S + c;
    const int64_t k0 = r0 + 1;
    double acc = 1.0 * col[r0 * lds];
    acc = wave_chain(
        acc, r1 - k0,
        [&](int64_t e, double &a, double &b) {
            a = v[k0 + e];
            b = col[(k0 + e) * lds];
        },
        stage[wave]);
    if ((threadIdx.x & 63) == 0) w[c] = acc * alpha;
}

// S[i][c] -= v[i] w[c] for i in [r0, r1), c in [c0, ncols) (v[r0] taken as 1).  With
// diag (the factorisation), column c0 then holds the stored reflector instead:
// diag[j] at row r0 and v below it (RealUnitaryDecomposition.updateData).
__global__ __launch_bounds__(UPD_COLS *UPD_ROWS) void hh_update_kernel(double *__restrict__ S, int64_t lds,
                                                                       int64_t c0, int64_t ncols, int64_t r0,
                                                                       int64_t r1, const double *__restrict__ v,
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
        const double vi = i == r0 ? 1.0 : v[i];
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
// chain from 0; a wave per right-hand side c, which keeps its solution column
// contiguous in xs (n doubles at xs + c n) for its own later rows.  U is the
// upper part of T (ld ldt).
__global__ __launch_bounds__(64 * BW) void back_solve_kernel(const double *__restrict__ T, int64_t ldt, int64_t n,
                                                             const double *__restrict__ Y, int64_t ldy,
                                                             double *__restrict__ X, int64_t ldx, int64_t R,
                                                             double *xs) {
    __shared__ __align__(16) double stage[BW][128];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * BW + wave;
    if (c >= R) return;
    double *x = xs + c * n;
    double xi = Y[(n - 1) * ldy + c] / T[(n - 1) * ldt + (n - 1)];
    if (lane == 0) {
        x[n - 1] = xi;
        X[(n - 1) * ldx + c] = xi;
    }
    for (int64_t i = n - 2; i >= 0; --i) {
        const double *u = T + i * ldt + (i + 1);
        const double *xk = x + (i + 1);
        const double x_next = xi;        // X[i + 1], just computed: taken from the register
        const double sum = wave_chain(
            0.0, n - 1 - i,
            [&](int64_t e, double &a, double &b) {
                a = u[e];
                b = e == 0 ? x_next : xk[e];
            },
            stage[wave]);
        xi = (Y[i * ldy + c] - sum) / T[i * ldt + i];
        if (lane == 0) {
            x[i] = xi;
            X[i * ldx + c] = xi;
        }
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
        hh_dot_kernel<<<(unsigned)((cols + CW - 1) / CW), 64 * CW, 0, stream>>>(T, ldt, j, N, j, M, v, qf + j, w);
        hh_update_kernel<<<update_grid(cols, M - j), dim3(UPD_COLS, UPD_ROWS), 0, stream>>>(T, ldt, j, N, j, M, v,
                                                                                            qf + j, w, diag + j);
    }
    return hipGetLastError();
}

hipError_t launch_householder_q(const double *T, int64_t M, int64_t N, int64_t ldt, const double *qf, double *Q,
                                int64_t ldq, double *v, double *w, hipStream_t stream) {
    const int64_t mn = M < N ? M : N;
    if (M * mn > 0) eye_kernel<<<(unsigned)((M * mn + 255) / 256), 256, 0, stream>>>(Q, M, mn, ldq);
    for (int64_t j = mn - 1; j >= 0; --j) {
        const int64_t cols = mn - j;
        hh_gather_kernel<<<(unsigned)((M - j + 255) / 256), 256, 0, stream>>>(T, ldt, M, j, v);
        hh_dot_kernel<<<(unsigned)((cols + CW - 1) / CW), 64 * CW, 0, stream>>>(Q, ldq, j, mn, j, M, v, qf + j, w);
        hh_update_kernel<<<update_grid(cols, M - j), dim3(UPD_COLS, UPD_ROWS), 0, stream>>>(Q, ldq, j, mn, j, M, v,
                                                                                            qf + j, w, nullptr);
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
                             int64_t ldx, int64_t R, double *xs, double *det, hipStream_t stream) {
    if (R > 0) back_solve_kernel<<<(unsigned)((R + BW - 1) / BW), 64 * BW, 0, stream>>>(T, ldt, n, Y, ldy, X, ldx, R, xs);
    det_kernel<<<1, 64, 0, stream>>>(T, ldt, n, R, det);
    return hipGetLastError();
}

}  // namespace f4j
