// Split-precision GEMM plan and launch interface (kernels in split_gemm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace f4j {

constexpr int SPLIT_MAX_SLICES = 8;
constexpr int SPLIT_MAX_SEGS = SPLIT_MAX_SLICES * (SPLIT_MAX_SLICES + 1) / 2;

// Number carrier of the split forms: BF16 or FP16 digits, or INT8 residues
// modulo coprime moduli (the modular form; batched plans only here).
enum SplitCarrier { SPLIT_BF16 = 0, SPLIT_FP16 = 1, SPLIT_INT8 = 2 };

// One launch of the slice-product GEMM: the digit pairs (s, t) whose levels
// L = s + t lie in [Llo, Lhi], summed exactly in FP64 with weights
// 2^shift = 2^(R (Lhi - L)); the sum is scaled by 2^(eA + eB + base_exp).
struct SplitGroup {
    int Llo, Lhi;
    int base_exp, nseg;
    int seg_a[SPLIT_MAX_SEGS], seg_b[SPLIT_MAX_SEGS], seg_shift[SPLIT_MAX_SEGS];
};

struct SplitPlan {
    int S = 0, carrier = SPLIT_BF16;
    int64_t Mp = 0, Np = 0, Kp = 0;     // padded sizes of the slice arrays (256 rows, 128 columns, 256 digits)
    size_t off_slices_a = 0, off_slices_b = 0, off_exp_a = 0, off_exp_b = 0, off_err = 0, off_residual = 0, bytes = 0;
    size_t off_flag = 0;                 // int: set by the exponent kernels when an operand holds NaN or Inf
    std::vector<SplitGroup> groups;     // in launch order (smallest weights first)
    int products() const { return S * (S + 1) / 2; }
};

struct SliceGemmArgs {
    const void *A, *B;                  // digit slices [S][Xp x Kp], fragment-major (slice_frag.h)
    int64_t Mp, Np, Kp, M, N;
    const int *eA, *eB;                 // row / column exponents
    double *C;
    int64_t ldc;
    int beta;                           // 0: C = result, 1: C += result
    int base_exp;
    int nseg;
    int seg_a[SPLIT_MAX_SEGS], seg_b[SPLIT_MAX_SEGS], seg_shift[SPLIT_MAX_SEGS];
};

int split_radix(int carrier);
int split_flush(int carrier);

// Level groups for S digits on `carrier` with padded inner size Kp, smallest
// weights first: each group's FP64 sum stays below 2^52, hence exact.
std::vector<SplitGroup> split_groups(int S, int carrier, int64_t Kp);

// Sizes, workspace layout and launch groups for C (M x N) = op(A) op(B),
// with room for a residual check of `check_rows` rows.
SplitPlan plan_split(int64_t M, int64_t N, int64_t K, int slices, int carrier, int check_rows);
size_t split_residual_bytes(int64_t N, int64_t K, int rows);

// Enqueue exponents, digit slices and slice products; C is overwritten.
// `ws` must hold plan.bytes bytes.  op codes: 0 = N, 1 = T.
// ev_begin / ev_end (optional) are recorded on `stream` around the slice products.
hipError_t launch_split_gemm(const SplitPlan &pl, int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A,
                             int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc, void *ws,
                             hipStream_t stream,
                             hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);

// Enqueue the FP64 residual check of `rows` evenly spaced rows of C; the
// maximum relative error (as double bits) is left at ws + plan.off_err.
hipError_t launch_split_residual(const SplitPlan &pl, int opA, int opB, int64_t M, int64_t N, int64_t K,
                                 const double *A, int64_t lda, const double *B, int64_t ldb, const double *C,
                                 int64_t ldc, int rows, void *ws, hipStream_t stream);

// Row exponents of op(A) (eA[M]) and column exponents of op(B) (eB[N]): the
// smallest E with |x| < 2^E over the line, INT_MIN for an all-zero line.
// *nonfinite is cleared, then set to 1 if any entry is NaN or infinite.
hipError_t launch_exponents(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                            const double *B, int64_t ldb, int *eA, int *eB, int *nonfinite, hipStream_t stream);

// Complex form (interleaved data, op codes 0 = N, 1 = T, 2 = C): |z| = |re| + |im|,
// an entry's error is the larger of its two parts.
hipError_t launch_residual_check_cplx(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A,
                                      int64_t lda, const double *B, int64_t ldb, const double *C, int64_t ldc,
                                      int rows, double *part, unsigned long long *err, hipStream_t stream);

// FP64 residual check of `rows` evenly spaced rows of C against compensated
// dot products; max |C - AB| / (|A||B|) left as double bits in *err.
// `part` holds split_residual_bytes(N, K, rows) bytes.
hipError_t launch_residual_check(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                                 const double *B, int64_t ldb, const double *C, int64_t ldc, int rows, double *part,
                                 unsigned long long *err, hipStream_t stream);
// The check in two phases: the compensated sums of the sampled rows need only
// A and B (they may run beside the product); the comparison needs C.
hipError_t launch_residual_partial(bool cplx, int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A,
                                   int64_t lda, const double *B, int64_t ldb, int rows, double *part,
                                   hipStream_t stream);
hipError_t launch_residual_combine(bool cplx, int64_t M, int64_t N, int64_t K, const double *C, int64_t ldc, int rows,
                                   const double *part, unsigned long long *err, hipStream_t stream);

}  // namespace f4j
