// Internal declarations shared by the flag4j_amd translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace f4j {

// Operand access kinds in the "real-expanded" view of a product.
// A complex128 product C = op(A) op(B) (M x N, inner K) is computed as the
// real product  C~ (M x 2N) = A~ (M x 2K) * B~ (2K x 2N)  where
//   A~[m][2k+c]      = component c of op(A)(m,k)
//   B~[2k][2n+c]     = component c of op(B)(k,n)
//   B~[2k+1][2n]     = -Im op(B)(k,n),   B~[2k+1][2n+1] = Re op(B)(k,n)
// so that C~ is exactly the interleaved (re, im) row-major output.
enum OpKind : int {
    R_N = 0,  // real, not transposed
    R_T = 1,  // real, transposed
    Z_N = 2,  // complex, not transposed
    Z_T = 3,  // complex, transposed
    Z_C = 4,  // complex, conjugate transposed
};

struct GemmProblem {
    int64_t M, N, K;          // real-expanded sizes: for complex N and K are doubled
    const double *A;
    int64_t lda;              // leading dimension of A's storage, in doubles
    const double *B;
    int64_t ldb;
    double *C;
    int64_t ldc;
    int flags;                // bit 0: 16-byte loads allowed (used by the grouped kernel)
};

// One tile of one problem in a grouped launch.
struct GroupedItem {
    int prob;                 // index into the problem table
    int tm, tn;               // tile row / column
};

// Launch the FP64 MFMA GEMM for one problem (kernel-level, device pointers).
hipError_t launch_gemm_f64(OpKind ka, OpKind kb, const GemmProblem &p, hipStream_t stream, int tile_cfg);

// Optional condition on the items of a grouped launch: with err set, an item of
// problem p runs only if slot[p] >= 0 and err[slot[p]] > above (the bits of
// non-negative doubles order as the values do).  The batched split plans use it
// to recompute in FP64 exactly the products whose residual check failed.
struct GroupedGuard {
    const unsigned long long *err = nullptr;
    const int *slot = nullptr;
    unsigned long long above = 0;
};

// Tile configurations from F4J_DIRECT_BASE on are the direct-load grouped
// kernels (real operands only): F4J_DIRECT_BASE + 0..3 = one wave per 32 x 32,
// 64 x 32, 32 x 64, 64 x 64 tile, one-wave workgroups.
constexpr int F4J_DIRECT_BASE = 100;
constexpr int F4J_NUM_DIRECT = 4;

// Persistent grouped FP64 GEMM over a work-item list; workgroup g runs items
// [begin[g], begin[g+1]).  All problems share op kinds and tile config.
hipError_t launch_gemm_f64_grouped(OpKind ka, OpKind kb, const GemmProblem *probs, const GroupedItem *items,
                                   const int *begin, int wgs, int cfg, hipStream_t stream,
                                   GroupedGuard guard = GroupedGuard{});

// Flag4j-order (bitwise) GEMM; sizes and leading dimensions in element units.
hipError_t launch_gemm_strict(bool cplx, int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A,
                              int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc,
                              hipStream_t stream);

// Flag4j-order (bitwise) mixed GEMM, no transposition: mode 0 real A x complex B,
// mode 1 complex A x real B; C complex.  Sizes and leading dimensions in elements.
hipError_t launch_gemm_strict_mixed(int mode, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                                    const double *B, int64_t ldb, double *C, int64_t ldc, hipStream_t stream);

// Matrix-vector products in the reference's summation orders (gemv.hip):
// GEMV_BLOCKED  real, 64-wide k blocks summed from +0.0 and added in block order;
// GEMV_CHAIN    real, one k-ascending sum per row;
// GEMV_CPLX     complex, one k-ascending sum of complex products per row;
// GEMV_RX       real A, complex x: per component one k-ascending sum of x_k a_ik;
// GEMV_CX       complex A, real x: per component one k-ascending sum of a_ik x_k.
// y[i * incy] = sum_k A[i * lda + k] x[k * incx]; lda, incx, incy in units of
// each operand's elements (complex or real), conjx conjugates x (GEMV_CPLX only).
// Returns hipErrorInvalidValue where the kernels do not apply (see gemv_supported;
// outside the blocked form, complex A and complex x must also be 16-byte aligned).
enum GemvKind : int { GEMV_BLOCKED = 0, GEMV_CHAIN = 1, GEMV_CPLX = 2, GEMV_RX = 3, GEMV_CX = 4 };
bool gemv_supported(int kind, int64_t K, int64_t lda, int64_t incx);
bool gemv_applicable(int kind, int64_t K, const double *A, int64_t lda, const double *x,
                     int64_t incx);   // supported and aligned
hipError_t launch_gemv(int kind, int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx,
                       double *y, int64_t incy, bool conjx, hipStream_t stream);

// Complex x real products as one real product (mixed.hip): split A (complex
// M x K, lda in complex elements) into W = [Re A; Im A] (2 x M x K doubles), and
// merge V (2 x M x N doubles) into the interleaved complex C (ldc in complex elements).
hipError_t launch_split_planes(const double *A, int64_t lda, int64_t M, int64_t K, double *W, hipStream_t stream);
hipError_t launch_merge_planes(const double *V, int64_t M, int64_t N, double *C, int64_t ldc, hipStream_t stream);

// Fill rows x cols (leading dimension ld, in doubles) with zeros.
hipError_t launch_zero_fill(double *C, int64_t rows, int64_t cols, int64_t ld, hipStream_t stream);

// Least squares in the reference's arithmetic (lstsq.hip).  T (M x N, ld ldt)
// holds A on entry and, after launch_householder_qr, R above the diagonal and the
// reflectors below it; v (M), w (N) are work vectors, qf (min(M, N)) the reflector
// factors, diag (min(M, N)) R's diagonal.  launch_householder_q forms the reduced
// Q (M x min(M, N), ld ldq) from T and qf (v, w: work vectors as above).
// launch_back_solve solves R X = Y (R the upper n x n part of T; Y, X n x R; xs:
// R n doubles of work space) and writes det R to `det`.
hipError_t launch_householder_qr(double *T, int64_t M, int64_t N, int64_t ldt, double *v, double *w, double *qf,
                                 double *diag, hipStream_t stream);
hipError_t launch_householder_q(const double *T, int64_t M, int64_t N, int64_t ldt, const double *qf, double *Q,
                                int64_t ldq, double *v, double *w, hipStream_t stream);
hipError_t launch_transpose(const double *Q, int64_t M, int64_t N, int64_t ldq, double *Qh, int64_t ldh,
                            hipStream_t stream);
hipError_t launch_back_solve(const double *T, int64_t ldt, int64_t n, const double *Y, int64_t ldy, double *X,
                             int64_t ldx, int64_t R, double *xs, double *det, hipStream_t stream);

void set_error(const std::string &msg);

// Initialise the library on first use and make its device current on the
// calling thread (returns an f4j_status).
int runtime_ready();

// The stream a NULL stream argument stands for (the library's own
// non-blocking stream); valid after runtime_ready().
hipStream_t runtime_stream(void *stream);

}  // namespace f4j
