// Split-precision FP64 GEMM: FP64 operands sliced into BF16 or FP16 integer
// digits, slice products on the BF16/FP16 matrix cores with FP32
// accumulation, recombined exactly in FP64 (an Ozaki-style scheme).
//
// No reference counterpart: Flag4j is FP64-only.  This is an alternative
// arithmetic for Matrix.mult (M:arrays/dense/Matrix.java:638-640,
// M: = src/main/java/org/flag4j/) selected by precision F4J_PREC_SPLIT_*.
//
// Representation.  Row i of A is scaled by 2^-eA[i] with eA[i] the binary
// exponent of its largest entry (|a| < 2^eA), so every scaled entry x lies in
// (-1, 1).  x is written with signed digits in radix 2^R:
//     x = 2^-(R-1) (d_1 + 2^-R (d_2 + 2^-R (d_3 + ...)))
//     d_1 = rint(x 2^(R-1)),  r_1 = x 2^(R-1) - d_1,  d_s = rint(r_{s-1} 2^R), ...
// so |d_s| <= 2^(R-1) and the remainder after S digits is at most
// 2^-(R-1) 2^-(R(S-1)) / 2 in units of 2^eA.  All steps are exact in FP64.
//   BF16 (8 significant bits) holds every integer |d| <= 256: R = 9.
//   FP16 (11 significant bits) holds every integer |d| <= 2048: R = 10.
// Columns of B are treated alike with exponents eB[j].
//
// Exactness of the products.  A digit product is at most 2^(2R-2) (2^16 for
// BF16, 2^18 for FP16).  FP32 represents every integer up to 2^24, so an
// FP32 MFMA accumulator is exact over FLUSH = 2^(24-2R+2) consecutive k
// (256 for BF16, 64 for FP16).  Every FLUSH k the FP32 accumulator is added,
// times a power of two, to an FP64 accumulator; the host groups the digit
// pairs so that this sum stays below 2^53, hence exact.  The only
// approximation left is the truncation: the pairs (s, t) with s + t > S + 1
// and the digits beyond S are dropped.
//
// Levels.  Pair (s, t) carries weight 2^(eA+eB) 2^-(2R-2) 2^-R(L) with level
// L = s + t - 2.  A launch sums a group of levels [Llo, Lhi] as
//     acc = sum_L 2^(R (Lhi - L)) (sum over pairs at level L, over k)
// and its epilogue adds 2^(eA+eB-(2R-2)-R Lhi) acc to C in FP64.  Groups run
// from the smallest weights to the largest, so C is rounded once per group.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <type_traits>
#include <cmath>
#include <cstdint>
#include <vector>

#include "amdgcn.h"
#include "f4j_internal.h"
#include "gemm_tile.h"
#include "slice_frag.h"
#include "split_gemm.h"
#include "split_tile.h"

namespace f4j {
namespace {

using slice::digits;
using slice::f32x16;
using slice::mfma32;
using slice::u32x4;


// ------------------------------------------------------------------------
// Exponents: eX[x] = max over k of the frexp exponent of X(x, k), i.e. the
// smallest E with |X(x, k)| < 2^E for all k (INT_MIN for an all-zero line).
// ------------------------------------------------------------------------

__global__ __launch_bounds__(256) void fill_int_kernel(int *p, int64_t n, int v) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

__device__ inline int exp_of(double v) {
    if (v == 0.0) return INT_MIN;
    int e;
    (void)frexp(v, &e);
    return e;
}

// Lines contiguous in memory (stored row x of length K): one block per line.
// A NaN or infinite entry sets *nonfinite (the digit slices cannot carry it).
__global__ __launch_bounds__(256) void exp_lines_kernel(const double *X, int64_t ld, int64_t K, int *e,
                                                        int *nonfinite) {
    const int64_t x = blockIdx.x;
    const double *row = X + x * ld;
    int m = INT_MIN;
    bool bad = false;
    for (int64_t k = threadIdx.x; k < K; k += 256) {
        bad |= !__builtin_isfinite(row[k]);
        m = max(m, exp_of(row[k]));
    }
    if (bad) *nonfinite = 1;
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    __shared__ int red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) e[x] = max(max(red[0], red[1]), max(red[2], red[3]));
}

// Lines strided in memory (stored K x X, line x is column x): each thread
// scans a k-chunk of one column and merges with atomicMax.
__global__ __launch_bounds__(256) void exp_cols_kernel(const double *X, int64_t ld, int64_t K, int64_t ncols,
                                                       int64_t kchunk, int *e, int *nonfinite) {
    const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (x >= ncols) return;
    const int64_t k0 = (int64_t)blockIdx.y * kchunk, k1 = min(K, k0 + kchunk);
    int m = INT_MIN;
    bool bad = false;
    for (int64_t k = k0; k < k1; ++k) {
        const double v = X[k * ld + x];
        bad |= !__builtin_isfinite(v);
        m = max(m, exp_of(v));
    }
    if (bad) *nonfinite = 1;
    if (m != INT_MIN) atomicMax(e + x, m);
}

// ------------------------------------------------------------------------
// Digit split.  Output: S slices, slice s at out + s*Xp*Kp, zero outside the
// X x K matrix; each slice fragment-major (dfrag::frag_offset with k counted in
// bytes: a thread's 8 consecutive digits of one line, k a multiple of 8, are one
// MFMA lane's 16 bytes).
// ------------------------------------------------------------------------

template <class T>
__device__ inline T *digits_at(T *out, int s, int64_t Xp, int64_t Kp, int64_t x, int64_t k0) {
    return (T *)((char *)out + (int64_t)s * Xp * Kp * 2 + dfrag::frag_offset(x, 2 * k0, Kp / 16));
}

// Lines contiguous: thread handles 8 consecutive k of one line.
template <int R, int S, class T>
__global__ __launch_bounds__(256) void split_lines_kernel(const double *X, int64_t ld, int64_t nx, int64_t K,
                                                          const int *e, T *out, int64_t Xp, int64_t Kp) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t k8 = Kp / 8;
    if (g >= Xp * k8) return;
    const int64_t x = g / k8, k0 = (g % k8) * 8;
    T d[8][S];
    const int ex = x < nx ? e[x] : INT_MIN;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int64_t k = k0 + u;
        const double v = (x < nx && k < K) ? X[x * ld + k] : 0.0;
        digits<R, S>(v, ex, d[u]);
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        T w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = d[u][s];
        *(u32x4 *)digits_at(out, s, Xp, Kp, x, k0) = __builtin_bit_cast(u32x4, w);
    }
}

// Lines strided (stored K x X): a 64 x 64 (k, x) block is transposed through
// LDS so the slice writes stay k-contiguous.
template <int R, int S, class T>
__global__ __launch_bounds__(256) void split_cols_kernel(const double *X, int64_t ld, int64_t nx, int64_t K,
                                                         const int *e, T *out, int64_t Xp, int64_t Kp) {
    __shared__ double tile[64][65];
    const int64_t x0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int64_t k = k0 + r, x = x0 + tx;
        tile[tx][r] = (k < K && x < nx) ? X[k * ld + x] : 0.0;
    }
    __syncthreads();
    // thread -> (x = x0 + t/8 ... ) : 64 x-lines x 8 groups of 8 k = 512 items, 2 per thread
    for (int item = threadIdx.x; item < 512; item += 256) {
        const int xl = item >> 3, kg = (item & 7) * 8;
        const int64_t x = x0 + xl;
        if (x >= Xp) continue;
        const int ex = x < nx ? e[x] : INT_MIN;
        T d[8][S];
#pragma unroll
        for (int u = 0; u < 8; ++u) digits<R, S>(tile[xl][kg + u], ex, d[u]);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            T w[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) w[u] = d[u][s];
            *(u32x4 *)digits_at(out, s, Xp, Kp, x, k0 + kg) = __builtin_bit_cast(u32x4, w);
        }
    }
}

// ------------------------------------------------------------------------
// Slice-product GEMM: C (+)= 2^(eA[i]+eB[j]+base) * sum_seg 2^shift_seg *
// sum_k A_{sa}[i][k] B_{sb}[j][k], both slice arrays k-contiguous.
// v_mfma_f32_32x32x16_{bf16,f16}: A/B lane l holds row l%32, k = 8(l/32)+v;
// result item r of lane l is row 8(r/4) + 4(l/32) + r%4, column l%32.
// ------------------------------------------------------------------------

constexpr int SBM = dfrag::BM, SBN = dfrag::BN;

// Fragment-major digit-slice products (slice_frag.h): a block computes a
// 256 x 128 tile with 8 waves of 64 x 64 over the launch's digit pairs, as one
// stream of 128-byte k steps (tps per pair); both operands' fragments go from
// the slices straight into a three-stage LDS ring (144 KB) by direct-to-LDS
// loads, each fragment loaded once per block, one barrier per step.  The FP32 accumulators
// are exact over FLUSH steps (BF16: 4 = 256 digits, FP16: 1 = 64) and are added,
// times 2^shift of their pair, into FP64 ones at least every FLUSH steps and at
// the end of each pair (the sums are exact integers, so where the flushes fall
// does not change the result).  At the barrier of step t every wave has read
// stage t % STAGES and has its own loads of step t + 1 retired (those of the
// STAGES - 2 later steps may still be in flight); after it step t + STAGES goes
// into stage t % STAGES (past the last step the last one again, so every step
// issues the same loads and the wait counts stay exact).
template <class T, int FLUSH>
__global__ __launch_bounds__(512, 2) void slice_frag_kernel(SliceGemmArgs a) {
    __shared__ __attribute__((aligned(1024))) unsigned char lds[dfrag::STAGES * dfrag::STAGE];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform: keeps descriptors scalar
    int64_t tm, tn;
    tile_of(blockIdx.x, a.Mp / SBM, a.Np / SBN, tm, tn);
    const int64_t m0 = (int64_t)__builtin_amdgcn_readfirstlane((int)tm) * SBM;
    const int64_t n0 = (int64_t)__builtin_amdgcn_readfirstlane((int)tn) * SBN;

    f32x16 acc[2][2];
    double acc64[2][2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                acc[i][j][r] = 0.f;
                acc64[i][j][r] = 0.0;
            }

    const int64_t Kpb = 2 * a.Kp, KC = Kpb >> 5;
    const int tps = (int)(Kpb / dfrag::SK);          // steps per digit pair
    const int nt = a.nseg * tps;
    const bool is_a = dfrag::loads_a(wave);
    const unsigned char *op = (const unsigned char *)(is_a ? a.A : a.B);
    const int64_t slice_bytes = (is_a ? a.Mp : a.Np) * Kpb;
    const int64_t x0 = is_a ? m0 : n0;

    auto issue = [&](int slot, int t) {              // step t into stage slot % STAGES
        const int seg = t / tps, kt = t - seg * tps;
        const int sl = is_a ? a.seg_a[seg] : a.seg_b[seg];
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(op + (int64_t)sl * slice_bytes + dfrag::load_src(x0, wave, 0, kt, KC)), (short)0, 0x7FFFFFF0,
            0x00020000);
        auto loads = [&](auto count) {
#pragma unroll
            for (int p = 0; p < decltype(count)::value; ++p)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rs, (lds_void *)(lds + dfrag::load_lds(slot, wave, p)), 16, dfrag::load_voff(lane),
                    (int)(dfrag::load_src(x0, wave, p, kt, KC) - dfrag::load_src(x0, wave, 0, kt, KC)), 0, 0);
        };
        if (is_a)                                    // uniform: A waves load LOADS_A fragments, B waves LOADS_B
            loads(std::integral_constant<int, dfrag::LOADS_A>{});
        else
            loads(std::integral_constant<int, dfrag::LOADS_B>{});
    };
    // Wait until this wave's loads of the oldest outstanding step are in LDS (the
    // STAGES - 2 newer steps' loads may stay in flight).
    auto wait_oldest = [&]() {
        static_assert(dfrag::STAGES >= 2 && (dfrag::STAGES - 2) * dfrag::LOADS_A < 64, "vmcnt holds 0..63");
        if (is_a)
            F4J_WAIT_VMCNT((dfrag::STAGES - 2) * dfrag::LOADS_A);
        else
            F4J_WAIT_VMCNT((dfrag::STAGES - 2) * dfrag::LOADS_B);
    };
#pragma unroll
    for (int s = 0; s + 1 < dfrag::STAGES; ++s) issue(s, min(s, nt - 1));
    wait_oldest();
    __builtin_amdgcn_s_barrier();
    issue(dfrag::STAGES - 1, min(dfrag::STAGES - 1, nt - 1));
    u32x4 fa[2][2], fb[2][2];
    // The two waves of a SIMD (w and w + 4) flush half an interval apart, so one's
    // FP64 work runs beside the other's MFMAs; every interval stays <= FLUSH steps.
    int since = dfrag::loads_a(wave) ? 0 : FLUSH / 2;
    for (int t = 0; t < nt; ++t) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 2; ++i) fa[0][i] = *(const u32x4 *)(lds + dfrag::a_frag(t, wave, i, 0, lane));
#pragma unroll
        for (int j = 0; j < 2; ++j) fb[0][j] = *(const u32x4 *)(lds + dfrag::b_frag(t, wave, j, 0, lane));
#pragma unroll
        for (int ks = 0; ks < dfrag::KB; ++ks) {
            if (ks + 1 < dfrag::KB) {
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    fa[(ks + 1) & 1][i] = *(const u32x4 *)(lds + dfrag::a_frag(t, wave, i, ks + 1, lane));
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    fb[(ks + 1) & 1][j] = *(const u32x4 *)(lds + dfrag::b_frag(t, wave, j, ks + 1, lane));
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = mfma32<T>(fa[ks & 1][i], fb[ks & 1][j], acc[i][j]);
        }
        const int seg = t / tps;
        if (++since == FLUSH || t + 1 == (seg + 1) * tps) {   // uniform
            const double sc = __builtin_bit_cast(double, (uint64_t)(1023 + a.seg_shift[seg]) << 52);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        acc64[i][j][r] = fma((double)acc[i][j][r], sc, acc64[i][j][r]);
                        acc[i][j][r] = 0.f;
                    }
            since = 0;
        }
        __builtin_amdgcn_sched_barrier(0);
        wait_oldest();                               // this wave's loads of step t + 1 are in LDS
        __builtin_amdgcn_s_waitcnt(0xC07F);          // lgkmcnt(0): its reads of stage t % STAGES are done
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(t, min(t + dfrag::STAGES, nt - 1));    // step t + STAGES into stage t % STAGES
    }
    F4J_WAIT_VMCNT(0);                               // no LDS write in flight when the block ends

    // Epilogue: C = beta C + 2^(eA+eB+base) acc64.
    const int64_t rw = m0 + dfrag::wave_row(wave), cw = n0 + dfrag::wave_col(wave);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t col = cw + dfrag::acc_col(j, lane);
            if (col >= a.N) continue;
            const int eb = a.eB[col];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t row = rw + dfrag::acc_row(i, r, lane);
                if (row >= a.M) continue;
                const int ea = a.eA[row];
                double v = 0.0;
                if (ea != INT_MIN && eb != INT_MIN) v = ldexp(acc64[i][j][r], ea + eb + a.base_exp);
                double *c = a.C + row * a.ldc + col;
                *c = a.beta ? (*c + v) : v;
            }
        }
}

// ------------------------------------------------------------------------
// Residual check on sampled rows: max over j of |C_ij - (AB)_ij| / (|A||B|)_ij
// with (AB)_ij from a compensated dot product (Dot2: TwoProduct by FMA and
// TwoSum; Ogita, Rump and Oishi 2005), accurate to ~u^2, so the figure is the
// split result's own error.
// ------------------------------------------------------------------------

constexpr int RES_KCHUNK = 256;

constexpr int RES_ROWS = 4;                       // sampled rows per thread (one read of B serves all)

// Pass 1: block (x, y, z) forms compensated partial dot products of sampled
// rows RES_ROWS y .. RES_ROWS y + 3 over k-chunk z for 256 columns.
__global__ __launch_bounds__(256) void residual_partial_kernel(const double *A, int64_t lda, int opA, const double *B,
                                                               int64_t ldb, int opB, int64_t M, int64_t N, int64_t K,
                                                               int64_t row_step, int nrows, double *part) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= N) return;
    const int r0 = blockIdx.y * RES_ROWS;
    int64_t ir[RES_ROWS];
#pragma unroll
    for (int u = 0; u < RES_ROWS; ++u) ir[u] = min(M - 1, (int64_t)min(r0 + u, nrows - 1) * row_step);
    const int64_t k0 = (int64_t)blockIdx.z * RES_KCHUNK, k1 = min(K, k0 + RES_KCHUNK);
    double p[RES_ROWS], q[RES_ROWS], sa[RES_ROWS];
#pragma unroll
    for (int u = 0; u < RES_ROWS; ++u) p[u] = q[u] = sa[u] = 0.0;
    for (int64_t k = k0; k < k1; ++k) {
        const double y = opB ? B[j * ldb + k] : B[k * ldb + j];
#pragma unroll
        for (int u = 0; u < RES_ROWS; ++u) {
            const double x = opA ? A[k * lda + ir[u]] : A[ir[u] * lda + k];
            const double h = x * y;
            const double r = fma(x, y, -h);
            const double s = p[u] + h;
            const double z = s - p[u];
            q[u] += ((p[u] - (s - z)) + (h - z)) + r;
            p[u] = s;
            sa[u] = fma(fabs(x), fabs(y), sa[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < RES_ROWS; ++u) {
        if (r0 + u >= nrows) break;
        double *o = part + (((int64_t)(r0 + u) * gridDim.z + blockIdx.z) * N + j) * 3;
        o[0] = p[u];
        o[1] = q[u];
        o[2] = sa[u];
    }
}

// Pass 2: combine the chunk partials of (row y, column j) with TwoSum and
// merge |C - AB| / (|A||B|) into the maximum.
__global__ __launch_bounds__(256) void residual_combine_kernel(const double *C, int64_t ldc, int64_t M, int64_t N,
                                                               int64_t row_step, int nchunks, const double *part,
                                                               unsigned long long *max_err_bits) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = min(M - 1, (int64_t)blockIdx.y * row_step);
    if (j >= N) return;
    double p = 0.0, q = 0.0, sa = 0.0;
    for (int c = 0; c < nchunks; ++c) {
        const double *o = part + (((int64_t)blockIdx.y * nchunks + c) * N + j) * 3;
        const double h = o[0];
        const double s = p + h;
        const double z = s - p;
        q += ((p - (s - z)) + (h - z)) + o[1];
        p = s;
        sa += o[2];
    }
    const double c = C[i * ldc + j];
    const double err = sa > 0.0 ? fabs((c - p) - q) / sa : (c == 0.0 ? 0.0 : INFINITY);
    atomicMax(max_err_bits, __builtin_bit_cast(unsigned long long, err));
}

// Complex form of the check (interleaved data, op codes N = 0, T = 1, C = 2):
// |z| = |re| + |im|, the error of an entry is the larger of its two parts.
__device__ inline void cload(const double *X, int op, int64_t ld, int64_t r, int64_t c, double &re, double &im) {
    // element (r, c) of op(X)
    const double *p = op ? X + 2 * (c * ld + r) : X + 2 * (r * ld + c);
    re = p[0];
    im = op == 2 ? -p[1] : p[1];
}

__device__ inline void dot2_step(double &p, double &q, double x, double y) {
    const double h = x * y;
    const double r = fma(x, y, -h);
    const double s = p + h;
    const double z = s - p;
    q += ((p - (s - z)) + (h - z)) + r;
    p = s;
}

__global__ __launch_bounds__(256) void residual_partial_cplx_kernel(const double *A, int64_t lda, int opA,
                                                                    const double *B, int64_t ldb, int opB, int64_t M,
                                                                    int64_t N, int64_t K, int64_t row_step,
                                                                    double *part) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = min(M - 1, (int64_t)blockIdx.y * row_step);
    if (j >= N) return;
    const int64_t k0 = (int64_t)blockIdx.z * RES_KCHUNK, k1 = min(K, k0 + RES_KCHUNK);
    double rp = 0.0, rq = 0.0, ip = 0.0, iq = 0.0, sa = 0.0;
    for (int64_t k = k0; k < k1; ++k) {
        double ar, ai, br, bi;
        cload(A, opA, lda, i, k, ar, ai);
        cload(B, opB, ldb, k, j, br, bi);
        dot2_step(rp, rq, ar, br);
        dot2_step(rp, rq, -ai, bi);
        dot2_step(ip, iq, ar, bi);
        dot2_step(ip, iq, ai, br);
        sa = fma(fabs(ar) + fabs(ai), fabs(br) + fabs(bi), sa);
    }
    double *o = part + (((int64_t)blockIdx.y * gridDim.z + blockIdx.z) * N + j) * 5;
    o[0] = rp;
    o[1] = rq;
    o[2] = ip;
    o[3] = iq;
    o[4] = sa;
}

__global__ __launch_bounds__(256) void residual_combine_cplx_kernel(const double *C, int64_t ldc, int64_t M, int64_t N,
                                                                    int64_t row_step, int nchunks, const double *part,
                                                                    unsigned long long *max_err_bits) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = min(M - 1, (int64_t)blockIdx.y * row_step);
    if (j >= N) return;
    double rp = 0.0, rq = 0.0, ip = 0.0, iq = 0.0, sa = 0.0;
    for (int c = 0; c < nchunks; ++c) {
        const double *o = part + (((int64_t)blockIdx.y * nchunks + c) * N + j) * 5;
        dot2_step(rp, rq, o[0], 1.0);
        rq += o[1];
        dot2_step(ip, iq, o[2], 1.0);
        iq += o[3];
        sa += o[4];
    }
    const double *c = C + 2 * (i * ldc + j);
    const double e = fmax(fabs((c[0] - rp) - rq), fabs((c[1] - ip) - iq));
    const double err = sa > 0.0 ? e / sa : (e == 0.0 ? 0.0 : INFINITY);
    atomicMax(max_err_bits, __builtin_bit_cast(unsigned long long, err));
}

template <int R, int S, class T>
void launch_split_one(const double *X, int64_t ld, bool lines, int64_t nx, int64_t K, const int *e, T *out, int64_t Xp,
                      int64_t Kp, hipStream_t stream) {
    if (lines) {
        const int64_t items = Xp * (Kp / 8);
        hipLaunchKernelGGL((split_lines_kernel<R, S, T>), dim3((unsigned)((items + 255) / 256)), dim3(256), 0, stream, X,
                           ld, nx, K, e, out, Xp, Kp);
    } else {
        hipLaunchKernelGGL((split_cols_kernel<R, S, T>), dim3((unsigned)(Xp / 64), (unsigned)(Kp / 64)), dim3(256), 0,
                           stream, X, ld, nx, K, e, out, Xp, Kp);
    }
}

template <int R, class T>
void launch_split_s(int S, const double *X, int64_t ld, bool lines, int64_t nx, int64_t K, const int *e, T *out,
                    int64_t Xp, int64_t Kp, hipStream_t stream) {
    switch (S) {
        case 1: launch_split_one<R, 1>(X, ld, lines, nx, K, e, out, Xp, Kp, stream); break;
        case 2: launch_split_one<R, 2>(X, ld, lines, nx, K, e, out, Xp, Kp, stream); break;
        case 3: launch_split_one<R, 3>(X, ld, lines, nx, K, e, out, Xp, Kp, stream); break;
        case 4: launch_split_one<R, 4>(X, ld, lines, nx, K, e, out, Xp, Kp, stream); break;
        case 5: launch_split_one<R, 5>(X, ld, lines, nx, K, e, out, Xp, Kp, stream); break;
        case 6: launch_split_one<R, 6>(X, ld, lines, nx, K, e, out, Xp, Kp, stream); break;
        case 7: launch_split_one<R, 7>(X, ld, lines, nx, K, e, out, Xp, Kp, stream); break;
        default: launch_split_one<R, 8>(X, ld, lines, nx, K, e, out, Xp, Kp, stream); break;
    }
}

}  // namespace

int split_radix(int carrier) { return carrier == SPLIT_FP16 ? 10 : 9; }
int split_flush(int carrier) { return carrier == SPLIT_FP16 ? 64 : 256; }

std::vector<SplitGroup> split_groups(int S, int carrier, int64_t Kp) {
    // Group the triangular pairs (s + t <= S + 1) by level, from the highest
    // level (smallest weight) down, keeping each group's FP64 sum exact:
    // sum_L 2^(R (Lhi - L)) npairs(L) Kp 2^(2R-2) < 2^52.
    std::vector<SplitGroup> out;
    const int R = split_radix(carrier);
    const double lim = std::ldexp(1.0, 52);
    int L = S - 1;
    while (L >= 0) {
        SplitGroup g{};
        g.Lhi = L;
        g.Llo = L;
        double bound = (double)(L + 1) * (double)Kp * std::ldexp(1.0, 2 * R - 2);
        while (g.Llo > 0) {
            const int Ln = g.Llo - 1;
            const double nb = bound + (double)(Ln + 1) * (double)Kp * std::ldexp(1.0, 2 * R - 2 + R * (L - Ln));
            if (nb >= lim) break;
            bound = nb;
            g.Llo = Ln;
        }
        g.base_exp = -(2 * R - 2) - R * g.Lhi;
        for (int Lv = g.Lhi; Lv >= g.Llo; --Lv)
            for (int s = 0; s <= Lv; ++s) {
                const int t = Lv - s;
                if (s >= S || t >= S) continue;
                g.seg_a[g.nseg] = s;
                g.seg_b[g.nseg] = t;
                g.seg_shift[g.nseg] = R * (g.Lhi - Lv);
                ++g.nseg;
            }
        out.push_back(g);
        L = g.Llo - 1;
    }
    return out;
}

SplitPlan plan_split(int64_t M, int64_t N, int64_t K, int slices, int carrier, int check_rows) {
    SplitPlan pl;
    pl.S = slices;
    pl.carrier = carrier;
    pl.Mp = (M + SBM - 1) / SBM * SBM;
    pl.Np = (N + SBN - 1) / SBN * SBN;
    pl.Kp = (K + 255) / 256 * 256;
    const size_t slice_bytes = (size_t)pl.S * (size_t)(pl.Mp + pl.Np) * (size_t)pl.Kp * 2;
    pl.off_slices_a = 0;
    pl.off_slices_b = (size_t)pl.S * pl.Mp * pl.Kp * 2;
    pl.off_exp_a = (slice_bytes + 255) & ~size_t(255);
    pl.off_exp_b = pl.off_exp_a + ((size_t)M * 4 + 255) / 256 * 256;
    pl.off_err = pl.off_exp_b + ((size_t)N * 4 + 255) / 256 * 256;
    pl.off_flag = pl.off_err + 128;
    pl.off_residual = pl.off_err + 256;
    pl.bytes = pl.off_residual + split_residual_bytes(N, K, (int)std::min<int64_t>(check_rows, M));

    pl.groups = split_groups(pl.S, carrier, pl.Kp);
    return pl;
}

hipError_t launch_split_gemm(const SplitPlan &pl, int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A,
                             int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc, void *ws,
                             hipStream_t stream, hipEvent_t ev_begin, hipEvent_t ev_end) {
    char *w = (char *)ws;
    void *sa = w + pl.off_slices_a, *sb = w + pl.off_slices_b;
    int *ea = (int *)(w + pl.off_exp_a), *eb = (int *)(w + pl.off_exp_b);
    const int S = pl.S;
    const bool fp16 = pl.carrier == SPLIT_FP16;

    if (hipError_t e = launch_exponents(opA, opB, M, N, K, A, lda, B, ldb, ea, eb, (int *)(w + pl.off_flag), stream);
        e != hipSuccess)
        return e;

    // digit slices
    auto split = [&](const double *X, int64_t ld, bool lines, int64_t nx, int64_t Xp, const int *e, void *out) {
        if (fp16)
            launch_split_s<10, _Float16>(S, X, ld, lines, nx, K, e, (_Float16 *)out, Xp, pl.Kp, stream);
        else
            launch_split_s<9, __bf16>(S, X, ld, lines, nx, K, e, (__bf16 *)out, Xp, pl.Kp, stream);
    };
    split(A, lda, opA == 0, M, pl.Mp, ea, sa);
    split(B, ldb, opB == 1, N, pl.Np, eb, sb);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;

    // slice products, smallest weights first
    if (ev_begin) (void)hipEventRecord(ev_begin, stream);
    bool first = true;
    for (const SplitGroup &g : pl.groups) {
        SliceGemmArgs a{};
        a.A = sa;
        a.B = sb;
        a.Mp = pl.Mp;
        a.Np = pl.Np;
        a.Kp = pl.Kp;
        a.M = M;
        a.N = N;
        a.eA = ea;
        a.eB = eb;
        a.C = C;
        a.ldc = ldc;
        a.beta = first ? 0 : 1;
        a.base_exp = g.base_exp;
        a.nseg = g.nseg;
        for (int q = 0; q < g.nseg; ++q) {
            a.seg_a[q] = g.seg_a[q];
            a.seg_b[q] = g.seg_b[q];
            a.seg_shift[q] = g.seg_shift[q];
        }
        const unsigned blocks = (unsigned)((pl.Mp / SBM) * (pl.Np / SBN));
        // flush every 64 digits (FP16: 1 step) or 256 (BF16: 4 steps), split_flush()
        if (fp16)
            hipLaunchKernelGGL((slice_frag_kernel<_Float16, 64 / (dfrag::SK / 2)>), dim3(blocks), dim3(64 * dfrag::WAVES),
                               0, stream, a);
        else
            hipLaunchKernelGGL((slice_frag_kernel<__bf16, 256 / (dfrag::SK / 2)>), dim3(blocks),
                               dim3(64 * dfrag::WAVES), 0, stream, a);
        first = false;
    }
    if (ev_end) (void)hipEventRecord(ev_end, stream);
    return hipGetLastError();
}

hipError_t launch_exponents(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                            const double *B, int64_t ldb, int *ea, int *eb, int *nonfinite, hipStream_t stream) {
    hipLaunchKernelGGL(fill_int_kernel, dim3(64), dim3(256), 0, stream, ea, M, INT_MIN);
    hipLaunchKernelGGL(fill_int_kernel, dim3(64), dim3(256), 0, stream, eb, N, INT_MIN);
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, stream, nonfinite, 1, 0);
    const int64_t kchunk = 256;
    if (M > 0 && K > 0) {
        if (opA == 0)
            hipLaunchKernelGGL(exp_lines_kernel, dim3((unsigned)M), dim3(256), 0, stream, A, lda, K, ea, nonfinite);
        else
            hipLaunchKernelGGL(exp_cols_kernel, dim3((unsigned)((M + 255) / 256), (unsigned)((K + kchunk - 1) / kchunk)),
                               dim3(256), 0, stream, A, lda, K, M, kchunk, ea, nonfinite);
    }
    if (N > 0 && K > 0) {
        if (opB == 1)
            hipLaunchKernelGGL(exp_lines_kernel, dim3((unsigned)N), dim3(256), 0, stream, B, ldb, K, eb, nonfinite);
        else
            hipLaunchKernelGGL(exp_cols_kernel, dim3((unsigned)((N + 255) / 256), (unsigned)((K + kchunk - 1) / kchunk)),
                               dim3(256), 0, stream, B, ldb, K, N, kchunk, eb, nonfinite);
    }
    return hipGetLastError();
}

hipError_t launch_split_residual(const SplitPlan &pl, int opA, int opB, int64_t M, int64_t N, int64_t K,
                                 const double *A, int64_t lda, const double *B, int64_t ldb, const double *C,
                                 int64_t ldc, int rows, void *ws, hipStream_t stream) {
    return launch_residual_check(opA, opB, M, N, K, A, lda, B, ldb, C, ldc, rows,
                                 (double *)((char *)ws + pl.off_residual),
                                 (unsigned long long *)((char *)ws + pl.off_err), stream);
}

hipError_t launch_residual_partial(bool cplx, int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A,
                                   int64_t lda, const double *B, int64_t ldb, int rows, double *part,
                                   hipStream_t stream) {
    const int r = (int)std::min<int64_t>(rows, M);
    if (r <= 0) return hipSuccess;
    const int64_t step = r > 1 ? (M - 1) / (r - 1) : 0;
    const int nch = (int)((K + RES_KCHUNK - 1) / RES_KCHUNK);
    if (cplx)
        hipLaunchKernelGGL(residual_partial_cplx_kernel, dim3((unsigned)((N + 255) / 256), (unsigned)r, (unsigned)nch),
                           dim3(256), 0, stream, A, lda, opA, B, ldb, opB, M, N, K, step, part);
    else
        hipLaunchKernelGGL(residual_partial_kernel,
                           dim3((unsigned)((N + 255) / 256), (unsigned)((r + RES_ROWS - 1) / RES_ROWS), (unsigned)nch),
                           dim3(256), 0, stream, A, lda, opA, B, ldb, opB, M, N, K, step, r, part);
    return hipGetLastError();
}

hipError_t launch_residual_combine(bool cplx, int64_t M, int64_t N, int64_t K, const double *C, int64_t ldc, int rows,
                                   const double *part, unsigned long long *err, hipStream_t stream) {
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(256), 0, stream, (int *)err, 2, 0);
    const int r = (int)std::min<int64_t>(rows, M);
    if (r <= 0) return hipGetLastError();
    const int64_t step = r > 1 ? (M - 1) / (r - 1) : 0;
    const int nch = (int)((K + RES_KCHUNK - 1) / RES_KCHUNK);
    if (cplx)
        hipLaunchKernelGGL(residual_combine_cplx_kernel, dim3((unsigned)((N + 255) / 256), (unsigned)r), dim3(256), 0,
                           stream, C, ldc, M, N, step, nch, part, err);
    else
        hipLaunchKernelGGL(residual_combine_kernel, dim3((unsigned)((N + 255) / 256), (unsigned)r), dim3(256), 0,
                           stream, C, ldc, M, N, step, nch, part, err);
    return hipGetLastError();
}

hipError_t launch_residual_check(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                                 const double *B, int64_t ldb, const double *C, int64_t ldc, int rows, double *part,
                                 unsigned long long *err, hipStream_t stream) {
    const hipError_t e = launch_residual_partial(false, opA, opB, M, N, K, A, lda, B, ldb, rows, part, stream);
    if (e != hipSuccess) return e;
    return launch_residual_combine(false, M, N, K, C, ldc, rows, part, err, stream);
}

hipError_t launch_residual_check_cplx(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A,
                                      int64_t lda, const double *B, int64_t ldb, const double *C, int64_t ldc,
                                      int rows, double *part, unsigned long long *err, hipStream_t stream) {
    const hipError_t e = launch_residual_partial(true, opA, opB, M, N, K, A, lda, B, ldb, rows, part, stream);
    if (e != hipSuccess) return e;
    return launch_residual_combine(true, M, N, K, C, ldc, rows, part, err, stream);
}

size_t split_residual_bytes(int64_t N, int64_t K, int rows) {
    const int64_t nch = (K + RES_KCHUNK - 1) / RES_KCHUNK;
    return (size_t)rows * (size_t)nch * (size_t)N * 5 * sizeof(double);     // room for the complex check
}

}  // namespace f4j
