/*
 * CPU oracle for the dense matrix-multiply path: TEST INFRASTRUCTURE ONLY.
 * Used by tests/, the smoke check and the benchmark's CPU-baseline leg as the
 * checker; never linked into or called by the product library.
 *
 * Restates the arithmetic (loop nests, summation order, rounding) of Flag4j's
 * CPU kernels.  Citation prefix M: = src/main/java/org/flag4j/.
 * Build: -O2 -ffp-contract=off -fno-fast-math (see Makefile), so every product
 * and sum is individually rounded as in Java (no FMA contraction).
 */
#include "f4j_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Real kernels                                                             */
/* ------------------------------------------------------------------------ */

/* M:linalg/ops/dense/real/RealDenseMatMult.java:50-79 (standard): dot form,
 * sum starts at +0.0, k ascending, then dest(=+0.0) += sum. */
void f4o_real_standard(int M, int N, int K, const double *A, const double *B, double *C) {
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) {
            double sum = 0.0;
            for (int k = 0; k < K; k++) sum += A[(int64_t)i * K + k] * B[(int64_t)k * N + j];
            C[(int64_t)i * N + j] = 0.0 + sum;
        }
}

/* RealDenseMatMult.java:91-122 (reordered): i-k-j, dest[i,:] += a[i,k]*b[k,:]. */
void f4o_real_reordered(int M, int N, int K, const double *A, const double *B, double *C) {
    memset(C, 0, sizeof(double) * (size_t)M * N);
    for (int i = 0; i < M; i++) {
        double *c = C + (int64_t)i * N;
        for (int k = 0; k < K; k++) {
            const double a = A[(int64_t)i * K + k];
            const double *b = B + (int64_t)k * N;
            for (int j = 0; j < N; j++) c[j] += a * b[j];
        }
    }
}

/* Same loop as reordered, accumulating with fma(): NOT a Flag4j kernel; it
 * exists so tests can show that the golden vectors reject contraction. */
void f4o_real_reordered_fma(int M, int N, int K, const double *A, const double *B, double *C) {
    memset(C, 0, sizeof(double) * (size_t)M * N);
    for (int i = 0; i < M; i++)
        for (int k = 0; k < K; k++)
            for (int j = 0; j < N; j++)
                C[(int64_t)i * N + j] = fma(A[(int64_t)i * K + k], B[(int64_t)k * N + j], C[(int64_t)i * N + j]);
}

static void blocked_reordered_rows(int N, int K, const double *A, const double *B, double *C,
                                   int ii0, int ii1, int bs) {
    /* RealDenseMatMult.java:188-231 / :386-407: for ii, kk, jj blocks, then
     * i, k, j with dest[i, jj..jBound) += a[i,k] * b[k, jj..jBound). */
    for (int ii = ii0; ii < ii1; ii += bs) {
        const int iB = ii + bs < ii1 ? ii + bs : ii1;
        for (int kk = 0; kk < K; kk += bs) {
            const int kB = kk + bs < K ? kk + bs : K;
            for (int jj = 0; jj < N; jj += bs) {
                const int jB = jj + bs < N ? jj + bs : N;
                for (int i = ii; i < iB; i++) {
                    double *c = C + (int64_t)i * N;
                    const double *a = A + (int64_t)i * K;
                    for (int k = kk; k < kB; k++) {
                        const double av = a[k];
                        const double *b = B + (int64_t)k * N;
                        for (int j = jj; j < jB; j++) c[j] += av * b[j];
                    }
                }
            }
        }
    }
}

void f4o_real_blocked_reordered(int M, int N, int K, const double *A, const double *B, double *C, int bs) {
    memset(C, 0, sizeof(double) * (size_t)M * N);
    blocked_reordered_rows(N, K, A, B, C, 0, M, bs);
}

/* RealDenseMatMult.java:458-491 (blockedVector): per 64-k block a fresh sum
 * from +0.0, then dest[i] += sum.  Differs in rounding from the k-ordered
 * closed form once K > bs. */
void f4o_real_blocked_vector(int M, int K, const double *A, const double *x, double *y, int bs) {
    for (int i = 0; i < M; i++) y[i] = 0.0;
    for (int ii = 0; ii < M; ii += bs) {
        const int iB = ii + bs < M ? ii + bs : M;
        for (int kk = 0; kk < K; kk += bs) {
            const int kB = kk + bs < K ? kk + bs : K;
            for (int i = ii; i < iB; i++) {
                double sum = 0.0;
                for (int k = kk; k < kB; k++) sum += A[(int64_t)i * K + k] * x[k];
                y[i] += sum;
            }
        }
    }
}

/* M:linalg/ops/dense/real/RealDenseMatMultTranspose.java:53-82 (multTranspose):
 * sum = dest (=+0.0), then sum += a[i,k]*b[j,k], k ascending.  The blocked and
 * concurrent variants (:95-233) reload sum from dest across k-blocks, so they
 * produce the same bits. */
void f4o_real_mult_transpose(int M, int N, int K, const double *A, const double *Bt, double *C) {
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) {
            double sum = 0.0;
            const double *a = A + (int64_t)i * K, *b = Bt + (int64_t)j * K;
            for (int k = 0; k < K; k++) sum += a[k] * b[k];
            C[(int64_t)i * N + j] = sum;
        }
}

/* ------------------------------------------------------------------------ */
/* Thread pool restating M:concurrency/ThreadManager.java                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    int kind;                    /* 0 real reordered, 1 real blocked, 2 complex reordered,
                                    3 real standard, 4 complex standard */
    int M, N, K, bs;
    const double *A, *B;
    double *C;
    int row0, row1;
    int chunk;                   /* kind 0/2: rows per task; kind 1: block size */
    int ntasks;
    int next;                    /* next task index (guarded by mu) */
    pthread_mutex_t mu;
} pool_job;

static void real_reordered_rows(int N, int K, const double *A, const double *B, double *C, int r0, int r1) {
    /* RealDenseMatMult.java:284-310 (concurrentReordered) task body. */
    for (int i = r0; i < r1; i++) {
        double *c = C + (int64_t)i * N;
        memset(c, 0, sizeof(double) * (size_t)N);
        const double *a = A + (int64_t)i * K;
        for (int k = 0; k < K; k++) {
            const double av = a[k];
            const double *b = B + (int64_t)k * N;
            for (int j = 0; j < N; j++) c[j] += av * b[j];
        }
    }
}

static void cplx_reordered_rows(int N, int K, const double *A, const double *B, double *C, int r0, int r1) {
    /* M:linalg/ops/dense/semiring_ops/DenseSemiringMatMult.java:288-314 task
     * body on primitive doubles; Complex128.mult/add (M:numbers/Complex128.java
     * :336-341, :276-278): p = (ar*br - ai*bi, ar*bi + ai*br), dest += p. */
    for (int i = r0; i < r1; i++) {
        double *c = C + (int64_t)i * 2 * N;
        memset(c, 0, sizeof(double) * 2 * (size_t)N);
        const double *a = A + (int64_t)i * 2 * K;
        for (int k = 0; k < K; k++) {
            const double ar = a[2 * k], ai = a[2 * k + 1];
            const double *b = B + (int64_t)k * 2 * N;
            for (int j = 0; j < N; j++) {
                const double br = b[2 * j], bi = b[2 * j + 1];
                const double pr = ar * br - ai * bi;
                const double pi = ar * bi + ai * br;
                c[2 * j] += pr;
                c[2 * j + 1] += pi;
            }
        }
    }
}

static void real_standard_rows(int N, int K, const double *A, const double *B, double *C, int r0, int r1) {
    /* RealDenseMatMult.java:243-272 (concurrentStandard) task body: one dot product
     * per entry, B read with stride N. */
    for (int i = r0; i < r1; i++)
        for (int j = 0; j < N; j++) {
            double sum = 0.0;
            for (int k = 0; k < K; k++) sum += A[(int64_t)i * K + k] * B[(int64_t)k * N + j];
            C[(int64_t)i * N + j] = 0.0 + sum;
        }
}

static void cplx_standard_rows(int N, int K, const double *A, const double *B, double *C, int r0, int r1) {
    /* DenseSemiringMatMult.java:247-287 (concurrentStandard) task body on primitive
     * doubles: per entry sum = ZERO, sum = sum.add(a.mult(b)) for k ascending. */
    for (int i = r0; i < r1; i++)
        for (int j = 0; j < N; j++) {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < K; k++) {
                const double ar = A[2 * ((int64_t)i * K + k)], ai = A[2 * ((int64_t)i * K + k) + 1];
                const double br = B[2 * ((int64_t)k * N + j)], bi = B[2 * ((int64_t)k * N + j) + 1];
                const double pr = ar * br - ai * bi;
                const double pi = ar * bi + ai * br;
                sr = sr + pr;
                si = si + pi;
            }
            C[2 * ((int64_t)i * N + j)] = sr;
            C[2 * ((int64_t)i * N + j) + 1] = si;
        }
}

static void *pool_worker(void *arg) {
    pool_job *job = (pool_job *)arg;
    for (;;) {
        pthread_mutex_lock(&job->mu);
        const int t = job->next++;
        pthread_mutex_unlock(&job->mu);
        if (t >= job->ntasks) break;
        const int s = job->row0 + t * job->chunk;
        const int e = s + job->chunk < job->row1 ? s + job->chunk : job->row1;
        if (s >= e) continue;
        if (job->kind == 0)
            real_reordered_rows(job->N, job->K, job->A, job->B, job->C, s, e);
        else if (job->kind == 1) {
            for (int i = s; i < e; i++) memset(job->C + (int64_t)i * job->N, 0, sizeof(double) * (size_t)job->N);
            blocked_reordered_rows(job->N, job->K, job->A, job->B, job->C, s, e, job->bs);
        } else if (job->kind == 2)
            cplx_reordered_rows(job->N, job->K, job->A, job->B, job->C, s, e);
        else if (job->kind == 3)
            real_standard_rows(job->N, job->K, job->A, job->B, job->C, s, e);
        else
            cplx_standard_rows(job->N, job->K, job->A, job->B, job->C, s, e);
    }
    return NULL;
}

static void run_pool(pool_job *job, int threads) {
    if (threads < 1) threads = 1;
    pthread_mutex_init(&job->mu, NULL);
    job->next = 0;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, pool_worker, job);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    pthread_mutex_destroy(&job->mu);
}

/* ThreadManager.java:155-168 (concurrentOperation): T tasks of ceil(R/T) rows. */
void f4o_real_mt_reordered(int M, int N, int K, const double *A, const double *B, double *C,
                           int row0, int row1, int threads) {
    (void)M;
    pool_job job = {0};
    const int R = row1 - row0, T = threads < 1 ? 1 : threads;
    job.kind = 0; job.N = N; job.K = K; job.A = A; job.B = B; job.C = C;
    job.row0 = row0; job.row1 = row1;
    job.chunk = (R + T - 1) / T; if (job.chunk < 1) job.chunk = 1;
    job.ntasks = T;
    run_pool(&job, T);
}

/* ThreadManager.java:202-228 (concurrentBlockedOperation): one task per
 * bs-row block on a T-thread pool. */
void f4o_real_mt_blocked_reordered(int M, int N, int K, const double *A, const double *B, double *C,
                                   int row0, int row1, int threads, int bs) {
    (void)M;
    pool_job job = {0};
    job.kind = 1; job.N = N; job.K = K; job.A = A; job.B = B; job.C = C; job.bs = bs;
    job.row0 = row0; job.row1 = row1; job.chunk = bs;
    job.ntasks = (row1 - row0 + bs - 1) / bs;
    run_pool(&job, threads);
}

void f4o_cplx_mt_reordered(int M, int N, int K, const double *A, const double *B, double *C,
                           int row0, int row1, int threads) {
    (void)M;
    pool_job job = {0};
    const int R = row1 - row0, T = threads < 1 ? 1 : threads;
    job.kind = 2; job.N = N; job.K = K; job.A = A; job.B = B; job.C = C;
    job.row0 = row0; job.row1 = row1;
    job.chunk = (R + T - 1) / T; if (job.chunk < 1) job.chunk = 1;
    job.ntasks = T;
    run_pool(&job, T);
}

/* ThreadManager.java:155-168 partitioning of the concurrentStandard kernels
 * (RealDenseMatMult.java:243-272; DenseSemiringMatMult.java:247-287, complex). */
void f4o_real_mt_standard(int M, int N, int K, const double *A, const double *B, double *C,
                          int row0, int row1, int threads) {
    (void)M;
    pool_job job = {0};
    const int R = row1 - row0, T = threads < 1 ? 1 : threads;
    job.kind = 3; job.N = N; job.K = K; job.A = A; job.B = B; job.C = C;
    job.row0 = row0; job.row1 = row1;
    job.chunk = (R + T - 1) / T; if (job.chunk < 1) job.chunk = 1;
    job.ntasks = T;
    run_pool(&job, T);
}

void f4o_cplx_mt_standard(int M, int N, int K, const double *A, const double *B, double *C,
                          int row0, int row1, int threads) {
    (void)M;
    pool_job job = {0};
    const int R = row1 - row0, T = threads < 1 ? 1 : threads;
    job.kind = 4; job.N = N; job.K = K; job.A = A; job.B = B; job.C = C;
    job.row0 = row0; job.row1 = row1;
    job.chunk = (R + T - 1) / T; if (job.chunk < 1) job.chunk = 1;
    job.ntasks = T;
    run_pool(&job, T);
}

/* ------------------------------------------------------------------------ */
/* Complex GEMM with op(A), op(B)                                           */
/* ------------------------------------------------------------------------ */

static inline void cload(const double *X, int op, int rows_op, int cols_op, int r, int c, double *re, double *im) {
    /* Element (r,c) of op(X), where op(X) is rows_op x cols_op. */
    int64_t idx;
    if (op == 0) idx = (int64_t)r * cols_op + c;          /* X stored rows_op x cols_op */
    else         idx = (int64_t)c * rows_op + r;          /* X stored cols_op x rows_op */
    *re = X[2 * idx];
    *im = (op == 2) ? -X[2 * idx + 1] : X[2 * idx + 1];   /* conj: Complex128.java:513-516 */
}

/* DenseSemiringMatMult.java:54-84 (standard) with T = Complex128, and
 * DenseSemiringMatMultTranspose.java:54-82 (multTranspose) for opB = T.
 * The reference forms A^H*B as A.H().mult(B) (conjugation is exact, so
 * reading conj on the fly gives the same bits). */
void f4o_cplx_gemm(int opA, int opB, int M, int N, int K, const double *A, const double *B, double *C) {
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < K; k++) {
                double ar, ai, br, bi;
                cload(A, opA, M, K, i, k, &ar, &ai);
                cload(B, opB, K, N, k, j, &br, &bi);
                const double pr = ar * br - ai * bi;
                const double pi = ar * bi + ai * br;
                sr = sr + pr;
                si = si + pi;
            }
            C[2 * ((int64_t)i * N + j)] = sr;
            C[2 * ((int64_t)i * N + j) + 1] = si;
        }
}

/* Sampled entries of Flag4j's products, for checks of large products where the
 * whole oracle product would take minutes: entry s is (op(A) op(B))_{I[s],J[s]}
 * as every kernel ReDeMatMultDispatcher / RealDenseMatrixMultiplyDispatcher can
 * select computes it (RealDenseMatMult.java:50-79 standard, and the reordered /
 * blocked forms, which give the same bits; RealDenseMatMultTranspose.java:53-82):
 * sum = +0.0; sum += a_ik * b_kj for k ascending, each product and sum rounded.
 * opA / opB: 0 = N, 1 = T (A stored K x M, B stored N x K). */
void f4o_real_entries(int opA, int opB, int M, int N, int K, const double *A, const double *B, int ns,
                      const int64_t *I, const int64_t *J, double *out) {
    for (int s = 0; s < ns; s++) {
        const int64_t i = I[s], j = J[s];
        double sum = 0.0;
        for (int k = 0; k < K; k++) {
            const double a = opA ? A[(int64_t)k * M + i] : A[i * K + k];
            const double b = opB ? B[j * K + k] : B[(int64_t)k * N + j];
            sum += a * b;
        }
        out[s] = sum;
    }
}

/* The complex counterpart (DenseSemiringMatMult.java:54-84 and the variants that
 * reload sum from dest, Complex128.java:276-278, :336-341, :513-516): per k
 * p = a.mult(b) then sum = sum.add(p) from (+0, +0).  out holds (re, im) pairs. */
void f4o_cplx_entries(int opA, int opB, int M, int N, int K, const double *A, const double *B, int ns,
                      const int64_t *I, const int64_t *J, double *out) {
    for (int s = 0; s < ns; s++) {
        const int64_t i = I[s], j = J[s];
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < K; k++) {
            double ar, ai, br, bi;
            cload(A, opA, M, K, (int)i, k, &ar, &ai);
            cload(B, opB, K, N, k, (int)j, &br, &bi);
            const double pr = ar * br - ai * bi;
            const double pi = ar * bi + ai * br;
            sr = sr + pr;
            si = si + pi;
        }
        out[2 * s] = sr;
        out[2 * s + 1] = si;
    }
}

/* Mixed real / complex products: RealFieldDenseMatMult
 * (M:linalg/ops/dense/real_field_ops/RealFieldDenseMatMult.java:58-1090).
 * Every variant (standard, reordered, blocked, concurrent, vector) sums each
 * C_ij from Complex128.ZERO in k order, sum = sum.add(z.mult(r)) with
 * Complex128.mult(double) = (re r, im r) (M:numbers/Complex128.java:402-404):
 * each component is a k-ascending sum of (component) * (real entry).
 * mode 0: A real M x K, B complex K x N; mode 1: A complex, B real.  C complex. */
void f4o_mixed_gemm(int mode, int M, int N, int K, const double *A, const double *B, double *C) {
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < K; k++) {
                double zr, zi, r;
                if (mode == 0) {
                    r = A[(int64_t)i * K + k];
                    zr = B[2 * ((int64_t)k * N + j)];
                    zi = B[2 * ((int64_t)k * N + j) + 1];
                } else {
                    zr = A[2 * ((int64_t)i * K + k)];
                    zi = A[2 * ((int64_t)i * K + k) + 1];
                    r = B[(int64_t)k * N + j];
                }
                sr = sr + zr * r;
                si = si + zi * r;
            }
            C[2 * ((int64_t)i * N + j)] = sr;
            C[2 * ((int64_t)i * N + j) + 1] = si;
        }
}

/* ------------------------------------------------------------------------ */
/* Dispatcher decision trees                                                */
/* ------------------------------------------------------------------------ */

/* M:linalg/ops/dispatch/ReDeMatMultDispatcher.java:157-171 (small path) and
 * :183-237 (getFunc).  m = A.rows, n = A.cols (inner), k = B.cols. */
int f4o_select_real(int m, int n, int k, const f4o_thresholds *t) {
    const long long total = (long long)m * n + (long long)n * k;
    if (total < t->small_threshold) return F4O_K_SMALL_STANDARD;
    if (k == 1) {
        const double ar = (double)(m > n ? m : n) / (double)(m < n ? m : n);
        if (ar <= 4 * t->aspect_threshold) return m < t->sq_seq_vec ? F4O_K_BLK_VEC : F4O_K_MT_BLK_VEC;
        return F4O_K_BLK_VEC;
    }
    int mn = k < m ? k : m; if (n < mn) mn = n;
    int mx = k > m ? k : m; if (n > mx) mx = n;
    const double ar = (double)mx / (double)mn;
    if (ar <= t->aspect_threshold) {
        if (n < t->sq_mt_standard) return F4O_K_MT_STRD;
        if (n < t->sq_mt_reordered) return F4O_K_MT_REORD;
        return F4O_K_MT_BLK_REORD;
    }
    if (m == mx) return mn < t->min_dim_small ? F4O_K_STRD : F4O_K_MT_REORD;
    if (n == mx) {
        if (mn < t->min_dim_small) return F4O_K_STRD;
        return m < t->wide_mt_reordered ? F4O_K_MT_STRD : F4O_K_MT_REORD;
    }
    return mn < t->min_dim_small ? F4O_K_REORD : F4O_K_MT_REORD;
}

/* M:linalg/ops/dispatch/Cm128DeMatMultDispatcher.java:155-170, :182-226: same
 * tree; N = 1 always goes to the concurrent standard vector kernel. */
int f4o_select_cplx(int m, int n, int k, const f4o_thresholds *t) {
    const long long total = (long long)m * n + (long long)n * k;
    if (total < t->small_threshold) return F4O_K_SMALL_STANDARD;
    if (k == 1) return F4O_K_MT_STRD_VEC;
    f4o_thresholds u = *t;
    u.sq_seq_vec = 0;
    return f4o_select_real(m, n, k, &u);
}

/* M:linalg/ops/RealDenseMatrixMultiplyDispatcher.java:228-245 and
 * M:linalg/ops/dense/semiring_ops/DenseSemiringMatMultDispatcher.java:255-271. */
int f4o_select_transpose(int rows) {
    if (rows < 40) return F4O_T_MULT_T;
    if (rows < 55) return F4O_T_BLOCKED;
    if (rows < 1200) return F4O_T_CONCURRENT;
    return F4O_T_BLOCKED_CONCURRENT;
}

/* ------------------------------------------------------------------------ */
/* Accuracy oracle: Dot2 (Ogita, Rump, Oishi, "Accurate sum and dot product",  */
/* SIAM J. Sci. Comput. 26(6), 2005, Algorithm 5.3).                          */
/* ------------------------------------------------------------------------ */

typedef struct { double p, s; } dot2_acc;

static inline void two_sum(double a, double b, double *s, double *e) {
    *s = a + b;
    const double z = *s - a;
    *e = (a - (*s - z)) + (b - z);
}

static inline void dot2_add(dot2_acc *acc, double x, double y) {
    const double h = x * y;
    const double r = fma(x, y, -h);          /* TwoProduct via FMA: exact error of x*y */
    double q;
    two_sum(acc->p, h, &acc->p, &q);
    acc->s += q + r;
}

static inline void dot2_result(const dot2_acc *acc, double *hi, double *lo) {
    const double h = acc->p + acc->s;
    *hi = h;
    *lo = acc->s - (h - acc->p);
}

void f4o_real_dot2(int M, int N, int K, const double *A, const double *B, int opB,
                   int ns, const int64_t *I, const int64_t *J, double *hi, double *lo, double *absab) {
    (void)M;
    for (int s = 0; s < ns; s++) {
        const int64_t i = I[s], j = J[s];
        dot2_acc acc = {0.0, 0.0};
        double ab = 0.0;
        for (int k = 0; k < K; k++) {
            const double a = A[i * K + k];
            const double b = opB ? B[j * K + k] : B[(int64_t)k * N + j];
            dot2_add(&acc, a, b);
            ab += fabs(a) * fabs(b);
        }
        dot2_result(&acc, &hi[s], &lo[s]);
        absab[s] = ab;
    }
}

void f4o_cplx_dot2(int opA, int opB, int M, int N, int K, const double *A, const double *B,
                   int ns, const int64_t *I, const int64_t *J,
                   double *re_hi, double *re_lo, double *im_hi, double *im_lo, double *absab) {
    for (int s = 0; s < ns; s++) {
        const int i = (int)I[s], j = (int)J[s];
        dot2_acc accr = {0.0, 0.0}, acci = {0.0, 0.0};
        double ab = 0.0;
        for (int k = 0; k < K; k++) {
            double ar, ai, br, bi;
            cload(A, opA, M, K, i, k, &ar, &ai);
            cload(B, opB, K, N, k, j, &br, &bi);
            dot2_add(&accr, ar, br);
            dot2_add(&accr, -ai, bi);
            dot2_add(&acci, ar, bi);
            dot2_add(&acci, ai, br);
            ab += (fabs(ar) + fabs(ai)) * (fabs(br) + fabs(bi));
        }
        dot2_result(&accr, &re_hi[s], &re_lo[s]);
        dot2_result(&acci, &im_hi[s], &im_lo[s]);
        absab[s] = ab;
    }
}

/* ------------------------------------------------------------------------ */
/* java.util.Random (Java SE API specification of java.util.Random)         */
/* ------------------------------------------------------------------------ */

#define LCG_MULT 0x5DEECE66DULL
#define LCG_ADD  0xBULL
#define LCG_MASK ((1ULL << 48) - 1)

void f4o_lcg_init(f4o_lcg *r, int64_t seed) {
    r->seed = ((uint64_t)seed ^ LCG_MULT) & LCG_MASK;
    r->have_next_gaussian = 0;
}

static inline int32_t lcg_next(f4o_lcg *r, int bits) {
    r->seed = (r->seed * LCG_MULT + LCG_ADD) & LCG_MASK;
    return (int32_t)(int64_t)(r->seed >> (48 - bits));
}

int32_t f4o_lcg_next_int(f4o_lcg *r) { return lcg_next(r, 32); }

double f4o_lcg_next_double(f4o_lcg *r) {
    const int64_t hi = (int64_t)(uint32_t)lcg_next(r, 26) & ((1 << 26) - 1);
    const int64_t lo = (int64_t)(uint32_t)lcg_next(r, 27) & ((1 << 27) - 1);
    return (double)((hi << 27) + lo) * 0x1.0p-53;
}

/* Polar method as specified for nextGaussian (log/sqrt from libm, not
 * StrictMath: same distribution, possibly different last bits). */
double f4o_lcg_next_gaussian(f4o_lcg *r) {
    if (r->have_next_gaussian) {
        r->have_next_gaussian = 0;
        return r->next_gaussian;
    }
    double v1, v2, s;
    do {
        v1 = 2 * f4o_lcg_next_double(r) - 1;
        v2 = 2 * f4o_lcg_next_double(r) - 1;
        s = v1 * v1 + v2 * v2;
    } while (s >= 1 || s == 0);
    const double mul = sqrt(-2 * log(s) / s);
    r->next_gaussian = v2 * mul;
    r->have_next_gaussian = 1;
    return v1 * mul;
}

void f4o_lcg_fill_uniform(int64_t seed, int64_t n, double *out) {
    f4o_lcg r;
    f4o_lcg_init(&r, seed);
    for (int64_t i = 0; i < n; i++) out[i] = 2.0 * f4o_lcg_next_double(&r) - 1.0;
}

/* ------------------------------------------------------------------------ */
/* Least squares, RealLstsqSolver.solve(A, B): reduced Householder QR       */
/* (M:linalg/decompositions/unitary/RealUnitaryDecomposition.java,          */
/* UnitaryDecomposition.java reduce(); M:linalg/decompositions/qr/RealQR.java;*/
/* M:linalg/transformations/Householder.java:156-188), X = back-solve of     */
/* R X = Q^H B (M:linalg/solvers/lstsq/LstsqSolver.java:160-178,             */
/* M:linalg/solvers/exact/triangular/RealBackSolver.java).                   */
/* ------------------------------------------------------------------------ */

static double jmax(double a, double b) {      /* java.lang.Math.max (NaN wins) */
    if (a != a || b != b) return NAN;
    return a > b ? a : b;
}

/* Householder.leftMultReflector: src -= (alpha v)(v^T src) on rows
 * [r0, r1) and columns [c0, ncols), as w = v^T src, w *= alpha, then the
 * rank-one update, every product and sum rounded. */
static void left_mult_reflector(double *src, int ncols, const double *v, double alpha, int c0, int r0, int r1,
                                double *w) {
    const double v0 = v[r0];
    for (int i = c0; i < ncols; i++) w[i] = v0 * src[(int64_t)r0 * ncols + i];
    for (int k = r0 + 1; k < r1; k++) {
        const double rv = v[k];
        const double *row = src + (int64_t)k * ncols;
        for (int i = c0; i < ncols; i++) w[i] += rv * row[i];
    }
    for (int i = c0; i < ncols; i++) w[i] = w[i] * alpha;            /* RealOps.scalMult */
    for (int i = r0; i < r1; i++) {
        const double rv = v[i];
        double *row = src + (int64_t)i * ncols;
        for (int j = c0; j < ncols; j++) row[j] -= rv * w[j];
    }
}

/* Reduced QR of A (M x N): T holds R above the diagonal and the reflectors
 * below it, qf the reflector factors; Q (M x min(M, N)) is formed as getQ does. */
void f4o_qr(int M, int N, const double *A, double *T, double *qf, double *Q) {
    const int mn = M < N ? M : N, mx = M > N ? M : N;
    double *v = (double *)calloc((size_t)mx, sizeof(double));
    double *w = (double *)calloc((size_t)mx, sizeof(double));
    memcpy(T, A, sizeof(double) * (size_t)M * N);
    const double eps = 0x1.0p-53;                                  /* Flag4jConstants.EPS_F64 */
    for (int j = 0; j < mn; j++) {
        double maxAbs = 0.0;                                       /* findMaxAndInit */
        for (int i = j; i < M; i++) {
            v[i] = T[(int64_t)i * N + j];
            maxAbs = jmax(fabs(v[i]), maxAbs);
        }
        double norm = 0.0, factor;
        const int apply = maxAbs >= eps;
        if (!apply) {
            factor = 0.0;
        } else {                                                   /* computePhasedNorm */
            const double maxInv = 1.0 / maxAbs;
            for (int i = j; i < M; i++) {
                v[i] *= maxInv;
                const double sv = v[i];
                norm += sv * sv;
            }
            norm = sqrt(norm);
            if (v[j] < 0) norm = -norm;
            const double shift = v[j] + norm;
            factor = shift / norm;
            norm *= maxAbs;
            const double invShift = 1.0 / shift;
            v[j] = 1.0;
            for (int i = j + 1; i < M; i++) v[i] *= invShift;
        }
        qf[j] = factor;
        if (apply) {                                               /* updateData */
            left_mult_reflector(T, N, v, qf[j], j, j, M, w);
            T[(int64_t)j * N + j] = -norm;
            for (int i = j + 1; i < M; i++) T[(int64_t)i * N + j] = v[i];
        }
    }
    if (Q) {                                                       /* getQ, reduced: I(M, mn) */
        memset(Q, 0, sizeof(double) * (size_t)M * mn);
        for (int i = 0; i < mn; i++) Q[(int64_t)i * mn + i] = 1.0;
        for (int j = mn - 1; j >= 0; j--) {
            v[j] = 1.0;
            for (int i = j + 1; i < M; i++) v[i] = T[(int64_t)i * N + j];
            if (qf[j] != 0) left_mult_reflector(Q, mn, v, qf[j], j, j, M, w);
        }
    }
    free(v);
    free(w);
}

/* X (N x R) for A (M x N), B (M x R): 0 on success, 1 when the back solver
 * reports a singular R (|det R| <= EPS_F64 max(N, N) or NaN; det is the product
 * of R's diagonal, last entry first), 2 when M < N (R is not square). */
int f4o_lstsq(int M, int N, int R, const double *A, const double *B, double *X) {
    if (M < N) return 2;
    double *T = (double *)malloc(sizeof(double) * (size_t)M * N);
    double *qf = (double *)malloc(sizeof(double) * (size_t)N);
    double *Q = (double *)malloc(sizeof(double) * (size_t)M * N);
    double *Y = (double *)malloc(sizeof(double) * (size_t)N * (R > 0 ? R : 1));
    f4o_qr(M, N, A, T, qf, Q);
    /* Y = Q^H B: Matrix.mult through ReDeMatMultDispatcher (or the matrix-vector
     * dispatch for one column): k-ascending sums, except one column above the
     * small path, which takes the 64-block rule (RealDenseMatMult.java:458-491). */
    const int small = (int64_t)N * M + (int64_t)M * R < 2450;
    for (int i = 0; i < N; i++)
        for (int c = 0; c < R; c++) {
            if (R == 1 && !small) {
                double dest = 0.0;
                for (int kk = 0; kk < M; kk += 64) {
                    double sum = 0.0;
                    for (int k = kk; k < M && k < kk + 64; k++) sum += Q[(int64_t)k * N + i] * B[k];
                    dest += sum;
                }
                Y[i] = dest;
            } else {
                double sum = 0.0;
                for (int k = 0; k < M; k++) sum += Q[(int64_t)k * N + i] * B[(int64_t)k * R + c];
                Y[(int64_t)i * R + c] = 0.0 + sum;
            }
        }
    /* RealBackSolver.solve(U, B) with U = R (N x N, the upper part of T) */
    const int n = N;
    double det = T[(int64_t)(n - 1) * N + (n - 1)];
    for (int i = n - 2; i >= 0; i--) det *= T[(int64_t)i * N + i];
    for (int c = 0; c < R; c++) {
        X[(int64_t)(n - 1) * R + c] = Y[(int64_t)(n - 1) * R + c] / T[(int64_t)(n - 1) * N + (n - 1)];
        for (int i = n - 2; i >= 0; i--) {
            double sum = 0.0;
            for (int k = i + 1; k < n; k++) sum += T[(int64_t)i * N + k] * X[(int64_t)k * R + c];
            X[(int64_t)i * R + c] = (Y[(int64_t)i * R + c] - sum) / T[(int64_t)i * N + i];
        }
    }
    const double ad = fabs(det);
    const int singular = ad <= 0x1.0p-53 * (double)n || ad != ad;
    free(T);
    free(qf);
    free(Q);
    free(Y);
    return singular ? 1 : 0;
}
