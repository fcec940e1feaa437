/* Drives every function of the CPU oracle (oracle/f4j_oracle.c) on small
 * random inputs, for builds under AddressSanitizer + UndefinedBehaviorSanitizer
 * and ThreadSanitizer (the thread-pool kernels).  Each multithreaded kernel's
 * result is compared with the sequential kernel it must equal bit for bit, so a
 * race that changes a result also fails here.
 *
 * usage: oracle_sanitize [threads]   (prints OK on success) */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/f4j_oracle.h"

static uint64_t g_state = 12345;
static double rnd(void) {
    g_state = g_state * 6364136223846793005ULL + 1442695040888963407ULL;
    const double u = (double)(g_state >> 11) * 0x1.0p-53;
    return (2 * u - 1) * pow(10.0, (int)(g_state % 7) - 3);
}
static double *mat(int64_t n) {
    double *p = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) p[i] = rnd();
    return p;
}
static int fails = 0;
static void same(const double *a, const double *b, int64_t n, const char *what) {
    if (memcmp(a, b, sizeof(double) * (size_t)n) != 0) {
        printf("FAIL %s\n", what);
        fails++;
    }
}

int main(int argc, char **argv) {
    const int T = argc > 1 ? atoi(argv[1]) : 4;
    const int shapes[][3] = {{1, 1, 1}, {7, 5, 3}, {33, 17, 65}, {64, 64, 64}, {70, 129, 130}};
    for (size_t s = 0; s < sizeof(shapes) / sizeof(shapes[0]); s++) {
        const int M = shapes[s][0], N = shapes[s][1], K = shapes[s][2];
        double *A = mat((int64_t)M * K), *B = mat((int64_t)K * N), *Bt = mat((int64_t)N * K);
        double *C0 = mat((int64_t)M * N), *C1 = mat((int64_t)M * N);
        f4o_real_standard(M, N, K, A, B, C0);
        f4o_real_reordered(M, N, K, A, B, C1);
        same(C0, C1, (int64_t)M * N, "reordered");
        f4o_real_reordered_fma(M, N, K, A, B, C1);
        f4o_real_blocked_reordered(M, N, K, A, B, C1, 16);
        same(C0, C1, (int64_t)M * N, "blocked_reordered");
        f4o_real_mt_reordered(M, N, K, A, B, C1, 0, M, T);
        same(C0, C1, (int64_t)M * N, "mt_reordered");
        f4o_real_mt_blocked_reordered(M, N, K, A, B, C1, 0, M, T, 8);
        same(C0, C1, (int64_t)M * N, "mt_blocked_reordered");
        f4o_real_mt_standard(M, N, K, A, B, C1, 0, M, T);
        same(C0, C1, (int64_t)M * N, "mt_standard");
        f4o_real_mult_transpose(M, N, K, A, Bt, C1);
        double *y = mat(M), *y2 = mat(M);
        f4o_real_blocked_vector(M, K, A, B, y, 64);
        f4o_real_blocked_vector(M, K, A, B, y2, 64);
        same(y, y2, M, "blocked_vector");
        int64_t I[8], J[8];
        for (int i = 0; i < 8; i++) {
            I[i] = (int64_t)(i * 7) % M;
            J[i] = (int64_t)(i * 5) % N;
        }
        double hi[8], lo[8], ab[8], e[16];
        f4o_real_dot2(M, N, K, A, B, 0, 8, I, J, hi, lo, ab);
        f4o_real_dot2(M, N, K, A, Bt, 1, 8, I, J, hi, lo, ab);
        f4o_real_entries(0, 0, M, N, K, A, B, 8, I, J, e);
        for (int i = 0; i < 8; i++)
            if (memcmp(&e[i], &C0[I[i] * N + J[i]], sizeof(double))) {
                printf("FAIL entries\n");
                fails++;
            }
        /* complex: interleaved pairs */
        double *Az = mat(2 * (int64_t)M * K), *Bz = mat(2 * (int64_t)K * N);
        double *Z0 = mat(2 * (int64_t)M * N), *Z1 = mat(2 * (int64_t)M * N);
        f4o_cplx_gemm(0, 0, M, N, K, Az, Bz, Z0);
        f4o_cplx_mt_reordered(M, N, K, Az, Bz, Z1, 0, M, T);
        same(Z0, Z1, 2 * (int64_t)M * N, "cplx_mt_reordered");
        f4o_cplx_mt_standard(M, N, K, Az, Bz, Z1, 0, M, T);
        same(Z0, Z1, 2 * (int64_t)M * N, "cplx_mt_standard");
        double *Azt = mat(2 * (int64_t)K * M);
        f4o_cplx_gemm(2, 1, M, N, K, Azt, Bz, Z1);           /* A^H B^T shapes: A stored K x M, B stored N x K */
        double rh[8], rl[8], ih[8], il[8];
        f4o_cplx_dot2(0, 0, M, N, K, Az, Bz, 8, I, J, rh, rl, ih, il, ab);
        f4o_cplx_entries(0, 0, M, N, K, Az, Bz, 8, I, J, e);
        f4o_mixed_gemm(0, M, N, K, A, Bz, Z1);
        f4o_mixed_gemm(1, M, N, K, Az, B, Z1);
        free(A); free(B); free(Bt); free(C0); free(C1); free(y); free(y2);
        free(Az); free(Bz); free(Z0); free(Z1); free(Azt);
    }
    f4o_thresholds th = {2450, 2.0, 50, 2048, 25, 256, 10};
    int sel = 0;
    for (int m = 1; m < 3000; m += 97)
        for (int n = 1; n < 3000; n += 89) {
            sel += f4o_select_real(m, n, m / 2 + 1, &th);
            sel += f4o_select_cplx(m, n, 1, &th);
        }
    sel += f4o_select_transpose(1300);
    f4o_lcg rng;
    f4o_lcg_init(&rng, 42);
    double acc = 0;
    for (int i = 0; i < 1000; i++) acc += f4o_lcg_next_double(&rng) + f4o_lcg_next_gaussian(&rng) + f4o_lcg_next_int(&rng);
    double *u = mat(1000);
    f4o_lcg_fill_uniform(7, 1000, u);
    free(u);
    if (fails) return 1;
    printf("OK selections=%d lcg=%g\n", sel, acc);
    return 0;
}
