// Layout passes for complex x real products.
//
// CMatrix.mult(Matrix) (M:arrays/dense/CMatrix.java:878-881 -> MatrixMultiplyDispatcher
// M:linalg/ops/MatrixMultiplyDispatcher.java:201-260 -> RealFieldDenseMatMult
// M:linalg/ops/dense/real_field_ops/RealFieldDenseMatMult.java:577-1090) computes each
// component of C_ij as a sum over k of (component of A_ik) * B_kj: the real and
// imaginary planes of A each multiply B on their own.  The library runs that as one
// real product [Re A; Im A] (2M x K) * B (K x N) on the matrix cores: these kernels
// split A into the two stacked planes and merge the 2M x N result back into the
// interleaved complex C.  Both are single streaming passes (16 bytes per complex
// element read and written).
#include <algorithm>

#include "f4j_internal.h"

namespace f4j {
namespace {

// W[c][i][k] = component c of A[i][k]; W is 2 x M x K, contiguous.
__global__ __launch_bounds__(256) void split_planes_kernel(const double *A, int64_t lda, int64_t M, int64_t K,
                                                           double *W) {
    const int64_t n = M * K;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int64_t i = t / K, k = t - i * K;
        const double *a = A + 2 * (i * lda + k);      // 8-byte accesses: A need not be 16-byte aligned
        W[t] = a[0];
        W[n + t] = a[1];
    }
}

// C[i][j] = (V[0][i][j], V[1][i][j]); V is 2 x M x N, contiguous.
__global__ __launch_bounds__(256) void merge_planes_kernel(const double *V, int64_t M, int64_t N, double *C,
                                                           int64_t ldc) {
    const int64_t n = M * N;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int64_t i = t / N, j = t - i * N;
        double *c = C + 2 * (i * ldc + j);
        c[0] = V[t];
        c[1] = V[n + t];
    }
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 8192); }

}  // namespace

hipError_t launch_split_planes(const double *A, int64_t lda, int64_t M, int64_t K, double *W, hipStream_t stream) {
    if (M * K == 0) return hipSuccess;
    hipLaunchKernelGGL(split_planes_kernel, dim3(grid_for(M * K)), dim3(256), 0, stream, A, lda, M, K, W);
    return hipGetLastError();
}

hipError_t launch_merge_planes(const double *V, int64_t M, int64_t N, double *C, int64_t ldc, hipStream_t stream) {
    if (M * N == 0) return hipSuccess;
    hipLaunchKernelGGL(merge_planes_kernel, dim3(grid_for(M * N)), dim3(256), 0, stream, V, M, N, C, ldc);
    return hipGetLastError();
}

}  // namespace f4j
