/*
 * CPU oracle for the dense matrix-multiply path: TEST INFRASTRUCTURE ONLY.
 *
 * A C restatement of the arithmetic of Flag4j's CPU matmul kernels, used as
 * the checker for the HIP kernels and as the timed CPU baseline.  Nothing in
 * the product library links, loads or calls this code.
 *
 * Citation prefix M: = src/main/java/org/flag4j/ in the Flag4j reference.
 *
 * Parity status of each group (how it is pinned):
 *   - f4o_real_*        : bitwise against the reference's golden vectors
 *                         (MatrixMultiplyTests, MatrixMultTests) and the NN==NT
 *                         invariants (RealDenseMatMultTransposeTests).
 *   - f4o_cplx_*        : bitwise against CMatrixMultiplyTests, CMatrixMultTests
 *                         and the ComplexDenseMatMultTransposeTests invariants.
 *   - f4o_mixed_gemm    : bitwise against MatrixMultTests.matMultComplexTestCase,
 *                         CMatrixMultTests.matMultTestCase and the mixed
 *                         matrix-vector tests (MatrixVectorTests,
 *                         CMatrixMatVecMultTests); agrees with f4o_cplx_gemm on a
 *                         real factor promoted to complex (inputs without zeros).
 *   - f4o_real_blocked_vector : weakly pinned (K = 3) by the column-vector
 *                         golden vectors; its 64-k partial-sum rule is checked
 *                         against a brute-force restatement at K > 64.
 *   - f4o_select_*      : decision tables restated from the dispatchers; pinned
 *                         by RealDenseMatrixMultiplyDispatcherTests for the
 *                         transpose selector, and by a table test of the
 *                         source's branches for the others.
 *   - f4o_dot2_*        : accuracy oracle (compensated Dot2); pinned against
 *                         exact rational arithmetic on small inputs.
 *   - f4o_lcg_*         : java.util.Random semantics; pinned against a second,
 *                         independent (jump-ahead) implementation and known
 *                         outputs of java.util.Random.
 *
 * All arithmetic is IEEE-754 binary64 with no contraction: build with
 * -O2 -ffp-contract=off -fno-fast-math (the Makefile beside this file does).
 */
#ifndef F4J_ORACLE_H
#define F4J_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- real, C = A*B, row-major contiguous; A MxK, B KxN, C MxN ---- */
void f4o_real_standard(int M, int N, int K, const double *A, const double *B, double *C);
void f4o_real_reordered(int M, int N, int K, const double *A, const double *B, double *C);
void f4o_real_blocked_reordered(int M, int N, int K, const double *A, const double *B, double *C, int bs);
/* Same with FMA accumulation: used only by tests to show the golden vectors
 * discriminate contraction. */
void f4o_real_reordered_fma(int M, int N, int K, const double *A, const double *B, double *C);
/* N = 1 blocked vector kernel: per-64-k partial sums added into dest. */
void f4o_real_blocked_vector(int M, int K, const double *A, const double *x, double *y, int bs);
/* C = A*B^T with B stored N x K. */
void f4o_real_mult_transpose(int M, int N, int K, const double *A, const double *Bt, double *C);

/* Multi-threaded CPU baseline: rows [row0,row1) of C only (rows are
 * independent), T worker threads, partition as Flag4j's ThreadManager. */
void f4o_real_mt_reordered(int M, int N, int K, const double *A, const double *B, double *C,
                           int row0, int row1, int threads);
void f4o_real_mt_blocked_reordered(int M, int N, int K, const double *A, const double *B, double *C,
                                   int row0, int row1, int threads, int bs);

/* ---- complex128, interleaved (re, im) row-major ---- */
/* op codes: 0 = N, 1 = T, 2 = C (conjugate transpose).  op(A) is MxK, op(B) is KxN.
 * Storage: opA==N -> A is MxK; opA in {T,C} -> A is KxM.  Same for B (KxN / NxK). */
void f4o_cplx_gemm(int opA, int opB, int M, int N, int K, const double *A, const double *B, double *C);
/* Mixed products (RealFieldDenseMatMult.java:58-1090): mode 0 real A x complex B, mode 1 complex A x real B. */
void f4o_mixed_gemm(int mode, int M, int N, int K, const double *A, const double *B, double *C);
void f4o_cplx_mt_reordered(int M, int N, int K, const double *A, const double *B, double *C,
                           int row0, int row1, int threads);
/* concurrentStandard (RealDenseMatMult.java:243-272, DenseSemiringMatMult.java:247-287):
 * T tasks of ceil(R/T) rows, one k-ascending dot product per entry. */
void f4o_real_mt_standard(int M, int N, int K, const double *A, const double *B, double *C,
                          int row0, int row1, int threads);
void f4o_cplx_mt_standard(int M, int N, int K, const double *A, const double *B, double *C,
                          int row0, int row1, int threads);

/* ---- dispatcher decision trees ---- */
typedef struct {
    int    small_threshold;          /* smallThreshold */
    double aspect_threshold;         /* aspectThreshold */
    int    sq_mt_standard;           /* squareMtStandardThreshold */
    int    sq_mt_reordered;          /* squareMtReorderedThreshold */
    int    wide_mt_reordered;        /* wideMtReorderedThreshold */
    int    sq_seq_vec;               /* squareSequentialVecThreshold */
    int    min_dim_small;            /* minDimSmallThreshold */
} f4o_thresholds;

enum {
    F4O_K_SMALL_STANDARD = 0, F4O_K_STRD, F4O_K_REORD, F4O_K_BLK, F4O_K_BLK_REORD,
    F4O_K_MT_STRD, F4O_K_MT_REORD, F4O_K_MT_BLK, F4O_K_MT_BLK_REORD,
    F4O_K_BLK_VEC, F4O_K_MT_BLK_VEC, F4O_K_MT_STRD_VEC
};
int f4o_select_real(int m, int n, int k, const f4o_thresholds *t);
int f4o_select_cplx(int m, int n, int k, const f4o_thresholds *t);
enum { F4O_T_MULT_T = 0, F4O_T_BLOCKED, F4O_T_CONCURRENT, F4O_T_BLOCKED_CONCURRENT };
int f4o_select_transpose(int rows);

/* ---- accuracy oracle: compensated dot products (Dot2) ---- */
/* Sampled entries of the reference's products in its own arithmetic (k-ascending,
 * separately rounded): real opA/opB 0 = N, 1 = T; complex opA/opB as f4o_cplx_gemm,
 * out as (re, im) pairs. */
void f4o_real_entries(int opA, int opB, int M, int N, int K, const double *A, const double *B, int ns,
                      const int64_t *I, const int64_t *J, double *out);
void f4o_cplx_entries(int opA, int opB, int M, int N, int K, const double *A, const double *B, int ns,
                      const int64_t *I, const int64_t *J, double *out);
/* For each sample s, hi[s]+lo[s] approximates (op(A) op(B))_{I[s],J[s]} with
 * error ~ u^2 * K * (|A||B|)_ij; absab[s] = (|A||B|)_ij.  Real, row-major,
 * opB: 0 = N (B is KxN), 1 = T (B is NxK). */
void f4o_real_dot2(int M, int N, int K, const double *A, const double *B, int opB,
                   int ns, const int64_t *I, const int64_t *J, double *hi, double *lo, double *absab);
/* Complex: separate results for re and im; absab is (|A||B|) with |z| taken as
 * |re|+|im| componentwise bound.  opA/opB as f4o_cplx_gemm. */
void f4o_cplx_dot2(int opA, int opB, int M, int N, int K, const double *A, const double *B,
                   int ns, const int64_t *I, const int64_t *J,
                   double *re_hi, double *re_lo, double *im_hi, double *im_lo, double *absab);

/* ---- java.util.Random restatement ---- */
typedef struct { uint64_t seed; int have_next_gaussian; double next_gaussian; } f4o_lcg;
void     f4o_lcg_init(f4o_lcg *r, int64_t seed);
int32_t  f4o_lcg_next_int(f4o_lcg *r);
double   f4o_lcg_next_double(f4o_lcg *r);
double   f4o_lcg_next_gaussian(f4o_lcg *r);
/* Fill with 2r-1 (uniform [-1,1)). */
void     f4o_lcg_fill_uniform(int64_t seed, int64_t n, double *out);

/* Least squares (RealLstsqSolver.solve): reduced Householder QR of A (M x N)
 * into T (R above the diagonal, reflectors below) and qf (factors), with Q
 * (M x min(M, N)) when Q is not null; f4o_lstsq returns 0, 1 (singular R) or
 * 2 (M < N) and writes X (N x R) on 0. */
void f4o_qr(int M, int N, const double *A, double *T, double *qf, double *Q);
int  f4o_lstsq(int M, int N, int R, const double *A, const double *B, double *X);

#ifdef __cplusplus
}
#endif
#endif
