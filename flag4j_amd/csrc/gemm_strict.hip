// "Strict" GEMM: Flag4j's own arithmetic on the GPU, bit for bit.
//
// Every Flag4j mat-mat kernel the dispatchers can select for N >= 2 computes
// each C_ij as a k-ascending sum, starting from +0.0, of individually rounded
// products, with no fused multiply-add:
//   real:    RealDenseMatMult.standard/reordered/blockedReordered and their
//            concurrent forms (M:linalg/ops/dense/real/RealDenseMatMult.java:50-122,
//            :188-231, :243-310, :374-415), RealDenseMatMultTranspose.multTranspose
//            (.../RealDenseMatMultTranspose.java:53-82);
//   complex: DenseSemiringMatMult.standard (M:linalg/ops/dense/semiring_ops/
//            DenseSemiringMatMult.java:54-84) with Complex128.mult/add
//            (M:numbers/Complex128.java:336-341, :276-278): per k,
//            p = (ar*br - ai*bi, ar*bi + ai*br), sum += p componentwise.
// This kernel reproduces that order per output element with contraction
// disabled, so its results equal the reference's exactly.  It runs on the
// FP64 vector ALUs (one thread per C element) and is meant for the small
// products the dispatcher routes here, where bitwise agreement with the
// reference's CPU results matters more than throughput.
#include <algorithm>

#include "f4j_internal.h"

#pragma clang fp contract(off)

namespace f4j {
namespace {

struct StrictArgs {
    int64_t M, N, K;            // element units
    const double *A;
    int64_t lda;                // element units
    const double *B;
    int64_t ldb;
    double *C;
    int64_t ldc;
    int opA, opB;               // 0 N, 1 T, 2 C
};

__global__ __launch_bounds__(256) void strict_real_kernel(StrictArgs a) {
    const int64_t j = (int64_t)blockIdx.x * 16 + (threadIdx.x & 15);
    const int64_t i = (int64_t)blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= a.M || j >= a.N) return;
    double sum = 0.0;
    for (int64_t k = 0; k < a.K; ++k) {
        const double x = a.opA == 0 ? a.A[i * a.lda + k] : a.A[k * a.lda + i];
        const double y = a.opB == 0 ? a.B[k * a.ldb + j] : a.B[j * a.ldb + k];
        const double p = __dmul_rn(x, y);
        sum = __dadd_rn(sum, p);
    }
    a.C[i * a.ldc + j] = sum;
}

__global__ __launch_bounds__(256) void strict_cplx_kernel(StrictArgs a) {
    const int64_t j = (int64_t)blockIdx.x * 16 + (threadIdx.x & 15);
    const int64_t i = (int64_t)blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= a.M || j >= a.N) return;
    double sr = 0.0, si = 0.0;
    for (int64_t k = 0; k < a.K; ++k) {
        const int64_t ia = a.opA == 0 ? i * a.lda + k : k * a.lda + i;
        const int64_t ib = a.opB == 0 ? k * a.ldb + j : j * a.ldb + k;
        const double ar = a.A[2 * ia], ai = a.opA == 2 ? -a.A[2 * ia + 1] : a.A[2 * ia + 1];
        const double br = a.B[2 * ib], bi = a.opB == 2 ? -a.B[2 * ib + 1] : a.B[2 * ib + 1];
        const double pr = __dsub_rn(__dmul_rn(ar, br), __dmul_rn(ai, bi));
        const double pi = __dadd_rn(__dmul_rn(ar, bi), __dmul_rn(ai, br));
        sr = __dadd_rn(sr, pr);
        si = __dadd_rn(si, pi);
    }
    a.C[2 * (i * a.ldc + j)] = sr;
    a.C[2 * (i * a.ldc + j) + 1] = si;
}

// Mixed products (RealFieldDenseMatMult, M:linalg/ops/dense/real_field_ops/
// RealFieldDenseMatMult.java:58-1090): each component of C_ij is a k-ascending
// sum from +0.0 of (component of the complex factor) * (real factor), with
// Complex128.mult(double) = (re b, im b) (M:numbers/Complex128.java:402-404).
// mode 0: A real (M x K), B complex (K x N); mode 1: A complex, B real.  Both
// stored row-major, leading dimensions in elements, no transposition.
__global__ __launch_bounds__(256) void strict_mixed_kernel(StrictArgs a, int mode) {
    const int64_t j = (int64_t)blockIdx.x * 16 + (threadIdx.x & 15);
    const int64_t i = (int64_t)blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= a.M || j >= a.N) return;
    double sr = 0.0, si = 0.0;
    for (int64_t k = 0; k < a.K; ++k) {
        double re, im, r;
        if (mode == 0) {
            r = a.A[i * a.lda + k];
            re = a.B[2 * (k * a.ldb + j)];
            im = a.B[2 * (k * a.ldb + j) + 1];
        } else {
            re = a.A[2 * (i * a.lda + k)];
            im = a.A[2 * (i * a.lda + k) + 1];
            r = a.B[k * a.ldb + j];
        }
        sr = __dadd_rn(sr, __dmul_rn(re, r));
        si = __dadd_rn(si, __dmul_rn(im, r));
    }
    a.C[2 * (i * a.ldc + j)] = sr;
    a.C[2 * (i * a.ldc + j) + 1] = si;
}

}  // namespace

hipError_t launch_gemm_strict_mixed(int mode, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                                    const double *B, int64_t ldb, double *C, int64_t ldc, hipStream_t stream) {
    if (M == 0 || N == 0) return hipSuccess;
    StrictArgs a{M, N, K, A, lda, B, ldb, C, ldc, 0, 0};
    dim3 grid((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16));
    hipLaunchKernelGGL(strict_mixed_kernel, grid, dim3(256), 0, stream, a, mode);
    return hipGetLastError();
}

hipError_t launch_gemm_strict(bool cplx, int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A,
                              int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc,
                              hipStream_t stream) {
    if (M == 0 || N == 0) return hipSuccess;
    StrictArgs a{M, N, K, A, lda, B, ldb, C, ldc, opA, opB};
    dim3 grid((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16));
    if (cplx)
        hipLaunchKernelGGL(strict_cplx_kernel, grid, dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(strict_real_kernel, grid, dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace f4j

namespace f4j {
namespace {
__global__ __launch_bounds__(256) void zero_fill_kernel(double *C, int64_t rows, int64_t cols, int64_t ld) {
    const int64_t n = rows * cols;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        C[(t / cols) * ld + t % cols] = 0.0;
}
}  // namespace

// Kernel-based fill (not hipMemsetAsync), so it is also safe inside captured graphs.
hipError_t launch_zero_fill(double *C, int64_t rows, int64_t cols, int64_t ld, hipStream_t stream) {
    const int64_t n = rows * cols;
    if (n == 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, C, rows, cols, ld);
    return hipGetLastError();
}

}  // namespace f4j
