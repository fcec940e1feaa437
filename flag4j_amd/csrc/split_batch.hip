// Batched split-precision GEMM (see split_batch.h).
//
// No reference counterpart: Flag4j multiplies one pair of matrices at a time
// in FP64 (M:arrays/dense/Matrix.java:638-640, M:arrays/dense/CMatrix.java:
// 650-652; M: = src/main/java/org/flag4j/).  The arithmetic is the one of the
// single-product split form (split_gemm.hip): per-line power-of-two scaling,
// signed radix-2^R digits on BF16/FP16, exact FP32 MFMA accumulation flushed
// into FP64, level groups summed exactly, recombined in FP64.
//
// Jobs.  A real product is one job.  A complex product op(A) op(B) with
// op(A) = a + i b and op(B) = c + i d (b, d carry the conjugation sign) is,
// in the 3M (Gauss) form, three real jobs
//     T1 = a c,  T2 = b d,  T3 = (a + b)(c + d),
//     Re C = T1 - T2,  Im C = T3 - T1 - T2,
// and in the 4M form four: a c, b d, a d, b c.  A job's operands ("sides")
// are read from the complex storage as ca re + cb im, so a + b is formed on
// the fly (one FP64 rounding, 2^-53 relative, before the exact split).
//
// Kernels, each one launch for the whole batch:
//   bexp_kernel    exponent of every line of every side
//   bsplit_kernel  digit slices [S][Xp][Kp] of every side
//   bgemm_kernel   persistent: every 64 x 64 output tile of every product,
//                  in a longest-first schedule; a tile runs all jobs of its
//                  product and all their level groups in sequence and adds
//                  each into C with the job's coefficients, so no two
//                  workgroups write the same element
//   bres_kernel    compensated (Dot2) check of sampled rows, per product
//
// Modular form (carrier SPLIT_INT8, the arithmetic of crt_gemm.hip): lines are
// scaled to t-bit integers (t from each product's K, at most 58) and stored as
// residues modulo n coprime moduli <= 256 in the fragment-major layout of the
// INT8 MFMA (crt_layout.h).  bcrt_split_kernel writes the residues of every
// side; bcrt_kernel runs every 64 x 64 output tile of every product: per job
// and modulus an exact int32 product on v_mfma_i32_32x32x32_i8, reduced mod m
// and accumulated straight into the entry's CRT chunk sums (exact FP64
// integers), recombined after the last modulus and added into C.  The product
// residues never leave the registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <climits>
#include <cmath>
#include <map>
#include <queue>
#include <vector>

#include "amdgcn.h"
#include "../../include/flag4j_amd.h"
#include "crt_gemm.h"
#include "crt_layout.h"
#include "crt_residue.h"
#include "f4j_internal.h"
#include "split_batch.h"
#include "split_gemm.h"
#include "split_tile.h"
#include "xcd_schedule.h"

namespace f4j {
namespace {

// Batched products use a 64 x 64 tile (2 x 2 waves of 32 x 32): small
// problems waste less padding, and the FP64 accumulators take 32 registers.
using BTile = slice::Tile<1, 1>;

struct Side {                 // one operand of one real job
    const double *X;          // real data, or interleaved complex (re at +0, im at +1)
    int64_t sx, sk;           // doubles between consecutive lines / consecutive k
    double ca, cb;            // value = ca re + cb im
    int cplx, contig;         // contig: consecutive k are adjacent elements
    int64_t nx, K, Xp, Kp;    // lines, inner size, padded to the tile / the k step
    int64_t dig;              // offset (carrier elements) of the digit slices [S][Xp][Kp]
                              // (modular form: bytes of the residues [n][Xp][Kp], fragment-major)
    int64_t eoff;             // offset of the line exponents [Xp]
    int prob;                 // the product's slot in the per-product error array
    int t;                    // modular form: bits of the scaled integers
};

// Error-array value that marks a product whose operands hold NaN or Inf: the
// bits of a quiet NaN, above every residual error's bits (finite or +Inf), so
// the residual check's atomicMax keeps it.
constexpr unsigned long long NONFINITE_BITS = 0x7FF8000000000000ull;

struct Job {
    int sa, sb;               // sides of op(A) and op(B)
    int grp0, ngrp;           // level groups of this job's Kp
    double cre, cim;          // C += cre T (real part), cim T (imaginary part)
};

struct Prob {
    int64_t M, N, K;
    const double *A, *B;
    double *C;
    int64_t lda, ldb, ldc;
    int job0, njobs;
};

struct Item {                 // one output tile
    int prob, tm, tn, cost;
};

__device__ inline double side_value(const Side &s, int64_t x, int64_t k) {
    const double *p = s.X + x * s.sx + k * s.sk;
    return s.cplx ? fma(s.ca, p[0], s.cb * p[1]) : p[0];
}

__device__ inline int exp_of(double v) {
    if (v == 0.0) return INT_MIN;
    int e;
    (void)frexp(v, &e);
    return e;
}

// Work item (side, first line): contiguous sides take 4 lines per block (a
// wave per line, lanes along k); strided sides 64 lines (lanes along the
// lines, so each k reads 64 adjacent entries; the block's four waves take
// every fourth k and merge in LDS).  Lines past nx (padding) get INT_MIN.
// A NaN or infinite entry marks the side's product in err (NONFINITE_BITS).
__global__ __launch_bounds__(256) void bexp_kernel(const Side *__restrict__ sides, const int2 *__restrict__ work,
                                                   int *__restrict__ epool, unsigned long long *__restrict__ err) {
    const int2 w = work[blockIdx.x];
    const Side s = sides[w.x];
    bool bad = false;
    if (s.contig) {
        const int64_t x = w.y + (threadIdx.x >> 6);
        const int lane = threadIdx.x & 63;
        int m = INT_MIN;
        if (x < s.nx)
            for (int64_t k = lane; k < s.K; k += 64) {
                const double v = side_value(s, x, k);
                bad |= !__builtin_isfinite(v);
                m = max(m, exp_of(v));
            }
        for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
        if (lane == 0 && x < s.Xp) epool[s.eoff + x] = m;
    } else {
        __shared__ int red[4][64];
        const int xl = threadIdx.x & 63, ph = threadIdx.x >> 6;
        const int64_t x = w.y + xl;                  // < Xp: Xp is a multiple of 64
        int m = INT_MIN;
        if (x < s.nx) {
#pragma unroll 4
            for (int64_t k = ph; k < s.K; k += 4) {
                const double v = side_value(s, x, k);
                bad |= !__builtin_isfinite(v);
                m = max(m, exp_of(v));
            }
        }
        red[ph][xl] = m;
        __syncthreads();                             // uniform: the whole block takes this branch
        if (ph == 0) epool[s.eoff + x] = max(max(red[0][xl], red[1][xl]), max(red[2][xl], red[3][xl]));
    }
    if (bad) atomicMax(err + s.prob, NONFINITE_BITS);
}

// Work item (side, x0, k0): 32 lines x 64 k staged through LDS with reads
// along k (contiguous sides) or along the lines (strided sides); thread t then
// splits line t / 8, 8 consecutive k, and writes 16 bytes per digit slice.
template <int R, int S, class T>
__global__ __launch_bounds__(256) void bsplit_kernel(const Side *__restrict__ sides, const int4 *__restrict__ work,
                                                     const int *__restrict__ epool, T *__restrict__ dpool) {
    __shared__ double tile[32][65];
    const int4 w = work[blockIdx.x];
    const Side s = sides[w.x];
    const int64_t x0 = w.y, k0 = w.z;
    const int t = threadIdx.x;
    if (s.contig) {
        const int c = t & 63;
        for (int r = t >> 6; r < 32; r += 4) {
            const int64_t x = x0 + r, k = k0 + c;
            tile[r][c] = (x < s.nx && k < s.K) ? side_value(s, x, k) : 0.0;
        }
    } else {
        const int r = t & 31;
        for (int c = t >> 5; c < 64; c += 8) {
            const int64_t x = x0 + r, k = k0 + c;
            tile[r][c] = (x < s.nx && k < s.K) ? side_value(s, x, k) : 0.0;
        }
    }
    __syncthreads();
    const int r = t >> 3, kq = (t & 7) * 8;
    const int64_t x = x0 + r;
    const int e = epool[s.eoff + x];
    T d[8][S];
#pragma unroll
    for (int u = 0; u < 8; ++u) slice::digits<R, S>(tile[r][kq + u], e, d[u]);
    T *out = dpool + s.dig + x * s.Kp + k0 + kq;
#pragma unroll
    for (int q = 0; q < S; ++q) {
        T v8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v8[u] = d[u][q];
        *(slice::u32x4 *)(out + (int64_t)q * s.Xp * s.Kp) = __builtin_bit_cast(slice::u32x4, v8);
    }
}

struct GroupSegs {
    const SplitGroup *__restrict__ g;
    __device__ int a(int q) const { return g->seg_a[q]; }
    __device__ int b(int q) const { return g->seg_b[q]; }
    __device__ int shift(int q) const { return g->seg_shift[q]; }
};

// Persistent slice-product kernel: workgroup w runs items begin[w] .. begin[w+1].
template <class T, int FLUSH>
// Tables are read-only for the kernel's lifetime (__restrict__): their loads
// are uniform and go through the scalar cache into SGPRs, which leaves the
// kernel at 165 VGPRs: 3 waves per SIMD (LDS 3 x 36 KB per CU).
__global__ __launch_bounds__(256, 3) void bgemm_kernel(const Item *__restrict__ items, const int *__restrict__ begin,
                                                       const Prob *__restrict__ probs, const Job *__restrict__ jobs,
                                                       const Side *__restrict__ sides,
                                                       const SplitGroup *__restrict__ groups,
                                                       const T *__restrict__ dpool, const int *__restrict__ epool,
                                                       int cplx) {
    __shared__ __attribute__((aligned(16))) T lds[2 * BTile::STAGE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int it1 = begin[blockIdx.x + 1];
    for (int it = begin[blockIdx.x]; it < it1; ++it) {
        const Item item = items[it];
        const Prob &p = probs[item.prob];
        const int64_t M = p.M, N = p.N, ldc = p.ldc;
        double *C = p.C;
        const int64_t m0 = (int64_t)item.tm * BTile::BM, n0 = (int64_t)item.tn * BTile::BN;
        bool first = true;
        for (int jb = 0; jb < p.njobs; ++jb) {
            const Job jo = jobs[p.job0 + jb];
            const Side &sa = sides[jo.sa];
            const Side &sb = sides[jo.sb];
            const T *da = dpool + sa.dig, *db = dpool + sb.dig;
            const int *ea = epool + sa.eoff, *eb = epool + sb.eoff;
            const int64_t Mp = sa.Xp, Np = sb.Xp, Kp = sa.Kp;
            for (int gi = 0; gi < jo.ngrp; ++gi) {
                const SplitGroup *g = groups + jo.grp0 + gi;
                double acc64[1][1][16];
                slice::tile_loop<T, FLUSH, 1, 1>(da, db, Mp, Np, Kp, g->nseg, GroupSegs{g}, m0, n0, lds, acc64);
                const int be = g->base_exp;
#pragma unroll
                for (int i = 0; i < 1; ++i)
#pragma unroll
                    for (int j = 0; j < 1; ++j) {
                        const int64_t col = n0 + slice::acc_col(wn, j, lane);
                        if (col >= N) continue;
                        const int ebv = eb[col];
#pragma unroll
                        for (int r = 0; r < 16; ++r) {
                            const int64_t row = m0 + slice::acc_row(wm, i, r, lane);
                            if (row >= M) continue;
                            const int eav = ea[row];
                            double v = 0.0;
                            if (eav != INT_MIN && ebv != INT_MIN) v = ldexp(acc64[i][j][r], eav + ebv + be);
                            if (cplx) {
                                double *c = C + 2 * (row * ldc + col);
                                c[0] = first ? jo.cre * v : c[0] + jo.cre * v;
                                c[1] = first ? jo.cim * v : c[1] + jo.cim * v;
                            } else {
                                double *c = C + row * ldc + col;
                                *c = first ? v : *c + v;
                            }
                        }
                    }
                first = false;
            }
        }
    }
}

__device__ inline void dot2_add(double &s, double &c, double a, double b) {
    const double h = a * b;
    const double r = fma(a, b, -h);
    const double t = s + h;
    const double z = t - s;
    c += ((s - (t - z)) + (h - z)) + r;
    s = t;
}

__global__ __launch_bounds__(256) void fill_u64_kernel(unsigned long long *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0ull;
}

// Work item (problem, checked row, first column; w = rows checked in this
// problem): compensated dot products (Dot2; Ogita, Rump and Oishi 2005) of
// row i = y M / w against every column; max |C - AB| / (|A||B|) per problem,
// |z| = |re| + |im| for complex entries.
__global__ __launch_bounds__(256) void bres_kernel(const Prob *__restrict__ probs, const int4 *__restrict__ work,
                                                   int cplx, int opA, int opB, unsigned long long *__restrict__ err) {
    const int4 w = work[blockIdx.x];
    const Prob &p = probs[w.x];
    const int64_t i = (int64_t)w.y * p.M / w.w;
    const int64_t j = w.z + threadIdx.x;
    if (j >= p.N) return;
    double e = 0.0;
    if (!cplx) {
        double s = 0.0, c = 0.0, sa = 0.0;
#pragma unroll 8
        for (int64_t k = 0; k < p.K; ++k) {         // (unrolled: the loads of 8 k run ahead of the sums)
            const double a = opA ? p.A[k * p.lda + i] : p.A[i * p.lda + k];
            const double b = opB ? p.B[j * p.ldb + k] : p.B[k * p.ldb + j];
            dot2_add(s, c, a, b);
            sa = fma(fabs(a), fabs(b), sa);
        }
        const double cv = p.C[i * p.ldc + j];
        e = sa > 0.0 ? fabs((cv - s) - c) / sa : (cv == 0.0 ? 0.0 : INFINITY);
    } else {
        const double sga = opA == 2 ? -1.0 : 1.0, sgb = opB == 2 ? -1.0 : 1.0;
        double rs = 0.0, rc = 0.0, is = 0.0, ic = 0.0, sa = 0.0;
#pragma unroll 8
        for (int64_t k = 0; k < p.K; ++k) {
            const double *pa = opA ? p.A + 2 * (k * p.lda + i) : p.A + 2 * (i * p.lda + k);
            const double *pb = opB ? p.B + 2 * (j * p.ldb + k) : p.B + 2 * (k * p.ldb + j);
            const double ar = pa[0], ai = sga * pa[1], br = pb[0], bi = sgb * pb[1];
            dot2_add(rs, rc, ar, br);
            dot2_add(rs, rc, -ai, bi);
            dot2_add(is, ic, ar, bi);
            dot2_add(is, ic, ai, br);
            sa = fma(fabs(ar) + fabs(ai), fabs(br) + fabs(bi), sa);
        }
        const double *cv = p.C + 2 * (i * p.ldc + j);
        const double er = fabs((cv[0] - rs) - rc), ei = fabs((cv[1] - is) - ic);
        e = sa > 0.0 ? fmax(er, ei) / sa : ((cv[0] == 0.0 && cv[1] == 0.0) ? 0.0 : INFINITY);
    }
    atomicMax(err + w.x, __builtin_bit_cast(unsigned long long, e));
}

// ---- modular (INT8) form ----


// Work item (side, x0, k0, -): the 32-line x 32-k fragment of every modulus,
// one wave per item, four per block.  WIDE (some t > 58): residues by the
// exact reduction step for |v| up to 2^75 (crt::residue_wide).  Each wave stages its 32 x 32 values
// through LDS (reads along k for contiguous sides, along the lines for strided
// ones); lane L then takes line L % 32 and k 16 (L / 32) .. + 15 and writes its
// 16 bytes of the fragment of every modulus (crt::frag_offset).
template <bool WIDE>
__global__ __launch_bounds__(256) void bcrt_split_kernel(const Side *__restrict__ sides, const int4 *__restrict__ work,
                                                         int64_t nwork, const int *__restrict__ epool,
                                                         signed char *__restrict__ rpool, CrtTables tb) {
    __shared__ double tile[4][32][33];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t wi = (int64_t)blockIdx.x * 4 + wave;
    const bool active = wi < nwork;
    const int4 w = work[active ? wi : 0];
    const Side s = sides[w.x];
    const int64_t x0 = w.y, k0 = w.z;
    if (active) {
        if (s.contig) {
            const int c = lane & 31;
            for (int r = lane >> 5; r < 32; r += 2) {
                const int64_t x = x0 + r, k = k0 + c;
                tile[wave][r][c] = (x < s.nx && k < s.K) ? side_value(s, x, k) : 0.0;
            }
        } else {
            const int r = lane & 31;
            for (int c = lane >> 5; c < 32; c += 2) {
                const int64_t x = x0 + r, k = k0 + c;
                tile[wave][r][c] = (x < s.nx && k < s.K) ? side_value(s, x, k) : 0.0;
            }
        }
    }
    __syncthreads();
    if (!active) return;
    const int r = lane & 31, kb = 16 * (lane >> 5);
    const int64_t x = x0 + r;
    const int ex = epool[s.eoff + x];
    const int sh = crt::scaled_shift(ex, s.t);
    double vh[16], vl[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const double v = crt::scaled_int_sh(tile[wave][r][kb + u], sh);
        if (WIDE) {
            vh[u] = v;                             // kept whole: the wide step reduces it per modulus
            vl[u] = 0.0;
        } else {
            crt::residue_split(v, vh[u], vl[u]);
            vl[u] = crt::residue_bias(vl[u]);     // the biased form (crt_residue.h)
        }
    }
    int4 *dst = (int4 *)(rpool + s.dig + crt::frag_offset(x, k0 + kb, s.Kp >> 5));
    const int64_t lstride = s.Xp * s.Kp / 16;
    for (int l = 0; l < tb.n; ++l) {
        const double m = (double)tb.m[l], mi = tb.minv[l], c26 = tb.c26[l], cq = crt::residue_cq(mi);
        unsigned rs[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
            rs[u] = WIDE ? crt::residue_wide_biased(vh[u], m, mi, cq, c26)
                         : crt::residue_biased(vh[u], vl[u], m, mi, cq, c26);
        int4 v;
        v.x = (int)crt::pack4_biased(rs[0], rs[1], rs[2], rs[3]);
        v.y = (int)crt::pack4_biased(rs[4], rs[5], rs[6], rs[7]);
        v.z = (int)crt::pack4_biased(rs[8], rs[9], rs[10], rs[11]);
        v.w = (int)crt::pack4_biased(rs[12], rs[13], rs[14], rs[15]);
        dst[l * lstride] = v;
    }
}

// Persistent residue-product kernel with the reconstruction fused: workgroup
// w runs output tiles begin[w] .. begin[w+1] (64 x 64, 2 x 2 waves of 32 x 32).
// For each job of the tile's product and each modulus l, a wave forms the
// exact int32 product of its 32 x 32 block over k, reduces it mod m_l
// (crt::product_residue_d: r in [0, m_l]) and adds r w[l][p] into the entry's
// chunk sums S_p (exact: below 2^48 per term and 2^53 over 20 moduli).  After
// the last modulus each entry is recombined (crt_recombine), scaled by
// 2^(eA + eB - 2t) and added into C with the job's coefficients; no two
// workgroups write one element.
//
// Operands: a job is a run of n KC / 2 steps (modulus-major), each step two k
// blocks of the block's 2 A and 2 B line blocks: 8 fragments of 1 KB, each
// contiguous in the fragment-major residues and loaded once per block (wave w
// loads line block w % 2 of A (w < 2) or of B (w >= 2), both k blocks, by
// buffer loads that write LDS directly) into a ring of RING steps.  At step s
// every wave waits for its own loads of step s (vmcnt), the block meets at a
// barrier, and each wave reads its A and B fragments and runs two MFMAs (one
// accumulator per k block parity); then it loads step s + RING - 1 into the
// slot every wave finished reading before that barrier.  Every step issues
// the same loads (past the job's last step they repeat it) so the wait count
// stays exact.  Loading each fragment once per block, rather than once per
// wave that needs it, halves the kernel's fragment traffic.
constexpr int RING = 4;                       // steps in flight per block
constexpr int STEP = 8 * crt::FRAG;           // LDS bytes of one step: A lb 0, 1 then B lb 0, 1; k blocks 2s, 2s + 1

__global__ __launch_bounds__(256, 2) void bcrt_kernel(const Item *__restrict__ items, const int *__restrict__ begin,
                                                      const Prob *__restrict__ probs, const Job *__restrict__ jobs,
                                                      const Side *__restrict__ sides,
                                                      const signed char *__restrict__ rpool,
                                                      const int *__restrict__ epool, CrtTables tb, int cplx) {
    __shared__ __attribute__((aligned(1024))) signed char ring[RING * STEP];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    // fragment slots within a step: A line block i at 2 i + kb, B line block j at 4 + 2 j + kb
    const int ld_base = (wave < 2 ? 2 * wave : 4 + 2 * (wave - 2)) * crt::FRAG;   // this wave's two loads
    const int rd_a = 2 * (wm >> 5) * crt::FRAG, rd_b = (4 + 2 * (wn >> 5)) * crt::FRAG;
    const int it1 = begin[blockIdx.x + 1];
    for (int it = begin[blockIdx.x]; it < it1; ++it) {
        const Item item = items[it];
        const Prob &p = probs[item.prob];
        const int64_t M = p.M, N = p.N, ldc = p.ldc;
        double *C = p.C;
        const int64_t m0 = (int64_t)item.tm * BTile::BM, n0 = (int64_t)item.tn * BTile::BN;
        for (int jb = 0; jb < p.njobs; ++jb) {
            const Job jo = jobs[p.job0 + jb];
            const Side &sa = sides[jo.sa];
            const Side &sb = sides[jo.sb];
            const int KC = (int)(sa.Kp >> 5);     // k blocks per modulus (even: Kp is a multiple of 64)
            const int total = tb.n * (KC >> 1);   // steps
            // this wave's line block of modulus 0 (A for waves 0-1, B for 2-3); moduli follow
            const Side &so = wave < 2 ? sa : sb;
            const int64_t x0 = (wave < 2 ? m0 : n0) + 32 * (wave & 1);
            const signed char *cur = rpool + so.dig + (((x0 >> 5) * KC) << 10);
            const int64_t plane = so.Xp * so.Kp;
            int iks = 0, issued = 0;              // next step to load: k blocks 2 iks, 2 iks + 1
            auto issue = [&](int slot) {
                const __amdgpu_buffer_rsrc_t rs =
                    __builtin_amdgcn_make_buffer_rsrc((void *)cur, (short)0, 0x7FFFFFF0, 0x00020000);
#pragma unroll
                for (int kb = 0; kb < 2; ++kb)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void *)(ring + slot * STEP + ld_base + kb * crt::FRAG),
                                                             16, 16 * lane, (2 * iks + kb) << 10, 0, 0);
                if (++issued < total && ++iks == (KC >> 1)) {   // past the last step: repeat it
                    iks = 0;
                    cur += plane;
                }
            };
#pragma unroll
            for (int j = 0; j < RING - 1; ++j) issue(j);
            double S0[16], S1[16], S2[16], S3[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) S0[r] = S1[r] = S2[r] = S3[r] = 0.0;
            i32x16 acc0, acc1;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc0[r] = acc1[r] = 0;
            int cks = 0, cl = 0;                   // step being multiplied (within modulus cl)
            for (int st = 0; st < total; ++st) {
                const int slot = st & (RING - 1);
                F4J_WAIT_VMCNT(2 * (RING - 2));      // this wave's loads of step st are in LDS
                __builtin_amdgcn_s_barrier();        // ... and every wave's
                __builtin_amdgcn_sched_barrier(0);
                const signed char *sp = ring + slot * STEP + 16 * lane;
                const i32x4 fa0 = *(const i32x4 *)(sp + rd_a), fa1 = *(const i32x4 *)(sp + rd_a + crt::FRAG);
                const i32x4 fb0 = *(const i32x4 *)(sp + rd_b), fb1 = *(const i32x4 *)(sp + rd_b + crt::FRAG);
                acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa0, fb0, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa1, fb1, acc1, 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                issue((st + RING - 1) & (RING - 1)); // the slot every wave read before this step's barrier
                if (++cks == (KC >> 1)) {          // modulus cl done: fold its residues into S
                    const double m = (double)tb.m[cl], minv = tb.minv[cl];
                    const double w0 = tb.w[cl][0], w1 = tb.w[cl][1], w2 = tb.w[cl][2], w3 = tb.w[cl][3];
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const double rr = crt::product_residue_d(acc0[r] + acc1[r], m, minv);
                        S0[r] = fma(rr, w0, S0[r]);
                        S1[r] = fma(rr, w1, S1[r]);
                        S2[r] = fma(rr, w2, S2[r]);
                        S3[r] = fma(rr, w3, S3[r]);
                        acc0[r] = acc1[r] = 0;
                    }
                    cks = 0;
                    ++cl;
                }
            }
            F4J_WAIT_VMCNT(0);                     // the repeated loads have landed
            __syncthreads();                       // and every wave is done with the ring
            const int64_t col = n0 + wn + crt::acc_col(0, lane);
            if (col < N) {
                const int ebv = epool[sb.eoff + col];
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int64_t row = m0 + wm + crt::acc_row(0, r, lane);
                    if (row >= M) continue;
                    const int eav = epool[sa.eoff + row];
                    double v = 0.0;
                    if (eav != INT_MIN && ebv != INT_MIN)
                        v = ldexp(crt_recombine(S0[r], S1[r], S2[r], S3[r], tb), eav + ebv - sa.t - sb.t);
                    if (cplx) {
                        double *c = C + 2 * (row * ldc + col);
                        c[0] = jb == 0 ? jo.cre * v : c[0] + jo.cre * v;
                        c[1] = jb == 0 ? jo.cim * v : c[1] + jo.cim * v;
                    } else {
                        double *c = C + row * ldc + col;
                        *c = jb == 0 ? v : *c + v;
                    }
                }
            }
        }
    }
}

template <int R, class T>
void launch_bsplit(int S, dim3 grid, hipStream_t st, const Side *sides, const int4 *work, const int *epool, T *dpool) {
    switch (S) {
        case 1: hipLaunchKernelGGL((bsplit_kernel<R, 1, T>), grid, dim3(256), 0, st, sides, work, epool, dpool); break;
        case 2: hipLaunchKernelGGL((bsplit_kernel<R, 2, T>), grid, dim3(256), 0, st, sides, work, epool, dpool); break;
        case 3: hipLaunchKernelGGL((bsplit_kernel<R, 3, T>), grid, dim3(256), 0, st, sides, work, epool, dpool); break;
        case 4: hipLaunchKernelGGL((bsplit_kernel<R, 4, T>), grid, dim3(256), 0, st, sides, work, epool, dpool); break;
        case 5: hipLaunchKernelGGL((bsplit_kernel<R, 5, T>), grid, dim3(256), 0, st, sides, work, epool, dpool); break;
        case 6: hipLaunchKernelGGL((bsplit_kernel<R, 6, T>), grid, dim3(256), 0, st, sides, work, epool, dpool); break;
        case 7: hipLaunchKernelGGL((bsplit_kernel<R, 7, T>), grid, dim3(256), 0, st, sides, work, epool, dpool); break;
        default: hipLaunchKernelGGL((bsplit_kernel<R, 8, T>), grid, dim3(256), 0, st, sides, work, epool, dpool); break;
    }
}

}  // namespace

struct SplitBatch {
    bool cplx = false;
    int opA = 0, opB = 0, carrier = SPLIT_BF16, S = 0, check_rows = 0;   // S: digits, or moduli (INT8)
    CrtTables tb{};                        // modular form: reconstruction tables for S moduli
    bool wide = false;                     // modular form: some product has t > 58
    int64_t count = 0;
    int real_products = 0;
    int workers = 0;
    int64_t n_exp = 0, n_split = 0, n_items = 0, n_res = 0;
    std::vector<BatchProblem> zero_k;      // products with K = 0: C is zero-filled
    // device tables
    Side *d_sides = nullptr;
    Job *d_jobs = nullptr;
    Prob *d_probs = nullptr;
    SplitGroup *d_groups = nullptr;
    Item *d_items = nullptr;
    int *d_begin = nullptr;
    int2 *d_exp_work = nullptr;
    int4 *d_split_work = nullptr, *d_res_work = nullptr;
    void *d_digits = nullptr;
    int *d_exps = nullptr;
    unsigned long long *d_err = nullptr;
    std::vector<int> prob_index;           // problem -> index in d_probs (-1: not in the GEMM)
    size_t bytes = 0;
};

namespace {

template <class X>
hipError_t upload(X **dst, const std::vector<X> &v, size_t *bytes) {
    if (v.empty()) return hipSuccess;
    hipError_t e = hipMalloc((void **)dst, v.size() * sizeof(X));
    if (e != hipSuccess) return e;
    *bytes += v.size() * sizeof(X);
    return hipMemcpy(*dst, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice);
}

}  // namespace

SplitBatch *split_batch_create(bool cplx, int opA, int opB, int carrier, int slices, bool use3m, int check_rows,
                               const std::vector<BatchProblem> &probs, std::string *err, int *status) {
    auto bad = [&](int code, const std::string &m) -> SplitBatch * {
        *status = code;
        *err = m;
        return nullptr;
    };
    const bool int8 = carrier == SPLIT_INT8;
    int64_t kmax = 1;
    for (const BatchProblem &q : probs)
        if (q.M > 0 && q.N > 0) kmax = std::max(kmax, q.K);
    CrtTables tb{};
    if (int8) {
        // |residue| <= 128, so the int32 products are exact for K < 2^17
        if (kmax >= (int64_t(1) << 17)) return bad(F4J_ERR_UNSUPPORTED, "modular form: K must be below 131072");
        // Default: t >= 64 bits at the largest K.  Short products of entries spread
        // over many binades need bits beyond 53: their (|A||B|)_ij is a sum of few
        // terms, far below 2^(eA + eB) (checked on 10^+-4 spreads at K = 32..1024).
        if (slices == 0) slices = crt_moduli_for_bits(64, kmax);
        if (slices < 2 || slices > CRT_MAX_MODULI)
            return bad(F4J_ERR_INVALID_ARG, "moduli must be in 2.." + std::to_string(CRT_MAX_MODULI));
        if (!crt_tables(slices, kmax, &tb))
            return bad(F4J_ERR_INVALID_ARG, std::to_string(slices) + " moduli leave no bits for K = " + std::to_string(kmax));
    } else if (slices < 1 || slices > SPLIT_MAX_SLICES) {
        return bad(F4J_ERR_INVALID_ARG, "slices must be in 1..8");
    }
    const int w = cplx ? 2 : 1;
    const double sga = (opA == F4J_OP_C) ? -1.0 : 1.0, sgb = (opB == F4J_OP_C) ? -1.0 : 1.0;

    SplitBatch *b = new SplitBatch;
    b->cplx = cplx;
    b->opA = opA;
    b->opB = opB;
    b->carrier = carrier;
    b->S = slices;
    b->tb = tb;
    b->check_rows = check_rows;
    b->count = (int64_t)probs.size();
    b->prob_index.assign(probs.size(), -1);

    std::vector<Side> sides;
    std::vector<Job> jobs;
    std::vector<Prob> dprobs;
    std::vector<SplitGroup> groups;
    std::map<int64_t, std::pair<int, int>> group_of_kp;     // Kp -> (first group, count)
    int64_t dig = 0, eoff = 0;

    // (ca_a, cb_a, ca_b, cb_b, cre, cim) per job of one product
    struct Form {
        double ca_a, cb_a, ca_b, cb_b, cre, cim;
    };
    std::vector<Form> forms;
    if (!cplx) forms = {{1, 0, 1, 0, 1, 0}};
    else if (use3m) forms = {{1, 0, 1, 0, 1, -1}, {0, sga, 0, sgb, -1, -1}, {1, sga, 1, sgb, 0, 1}};
    else forms = {{1, 0, 1, 0, 1, 0}, {0, sga, 0, sgb, -1, 0}, {1, 0, 0, sgb, 0, 1}, {0, sga, 1, 0, 0, 1}};

    for (size_t pi = 0; pi < probs.size(); ++pi) {
        const BatchProblem &q = probs[pi];
        if (q.M == 0 || q.N == 0) continue;
        if (q.K == 0) {
            b->zero_k.push_back(q);
            continue;
        }
        const int64_t Mp = (q.M + BTile::BM - 1) / BTile::BM * BTile::BM;
        const int64_t Np = (q.N + BTile::BN - 1) / BTile::BN * BTile::BN;
        const int64_t Kp = (q.K + slice::BK - 1) / slice::BK * slice::BK;   // 64: whole pairs of k blocks
        // modular form: t from this product's K (more bits for short products), at most
        // 75, the range of the residue steps (crt_residue.h)
        const int t = int8 ? std::min(75, crt_bits(slices, q.K)) : 0;
        b->wide |= t > 58;
        auto git = group_of_kp.find(Kp);
        if (int8) {
            git = group_of_kp.emplace(0, std::make_pair(0, 0)).first;     // no level groups
        } else if (git == group_of_kp.end()) {
            const std::vector<SplitGroup> g = split_groups(slices, carrier, Kp);
            git = group_of_kp.emplace(Kp, std::make_pair((int)groups.size(), (int)g.size())).first;
            groups.insert(groups.end(), g.begin(), g.end());
        }
        Prob pr{};
        pr.M = q.M;
        pr.N = q.N;
        pr.K = q.K;
        pr.A = q.A;
        pr.B = q.B;
        pr.C = q.C;
        pr.lda = q.lda;
        pr.ldb = q.ldb;
        pr.ldc = q.ldc;
        pr.job0 = (int)jobs.size();
        pr.njobs = (int)forms.size();
        for (const Form &f : forms) {
            Side a{};
            a.X = q.A;
            a.cplx = cplx;
            a.ca = f.ca_a;
            a.cb = f.cb_a;
            a.nx = q.M;
            a.K = q.K;
            a.Xp = Mp;
            a.Kp = Kp;
            a.contig = opA == F4J_OP_N;
            a.sx = (a.contig ? q.lda : 1) * w;
            a.sk = (a.contig ? 1 : q.lda) * w;
            Side c{};
            c.X = q.B;
            c.cplx = cplx;
            c.ca = f.ca_b;
            c.cb = f.cb_b;
            c.nx = q.N;
            c.K = q.K;
            c.Xp = Np;
            c.Kp = Kp;
            c.contig = opB != F4J_OP_N;
            c.sx = (c.contig ? q.ldb : 1) * w;
            c.sk = (c.contig ? 1 : q.ldb) * w;
            for (Side *s : {&a, &c}) {
                s->prob = (int)dprobs.size();
                s->t = t;
                s->dig = dig;
                dig += (int64_t)slices * s->Xp * s->Kp;
                s->eoff = eoff;
                eoff += s->Xp;
            }
            Job j{};
            j.sa = (int)sides.size();
            sides.push_back(a);
            j.sb = (int)sides.size();
            sides.push_back(c);
            j.grp0 = git->second.first;
            j.ngrp = git->second.second;
            j.cre = f.cre;
            j.cim = f.cim;
            jobs.push_back(j);
        }
        b->prob_index[pi] = (int)dprobs.size();
        dprobs.push_back(pr);
    }
    b->real_products = (int)jobs.size() * (int8 ? slices : slices * (slices + 1) / 2);

    // Work lists.
    std::vector<int2> exp_work;
    std::vector<int4> split_work;
    for (size_t si = 0; si < sides.size(); ++si) {
        const Side &s = sides[si];
        const int step = s.contig ? 4 : 64;
        for (int64_t x = 0; x < s.Xp; x += step) exp_work.push_back(make_int2((int)si, (int)x));
        const int64_t kstep = int8 ? 32 : slice::BK;
        for (int64_t x = 0; x < s.Xp; x += 32)
            for (int64_t k = 0; k < s.Kp; k += kstep) split_work.push_back(make_int4((int)si, (int)x, (int)k, 0));
    }
    std::vector<int4> res_work;
    if (check_rows > 0)
        for (size_t pi = 0; pi < dprobs.size(); ++pi) {
            const Prob &p = dprobs[pi];
            const int rows = (int)std::min<int64_t>(check_rows, p.M);
            for (int y = 0; y < rows; ++y)
                for (int64_t c = 0; c < p.N; c += 256) res_work.push_back(make_int4((int)pi, y, (int)c, rows));
        }

    // Output tiles, longest first onto the least-loaded workgroup.
    std::vector<Item> items;
    for (size_t pi = 0; pi < dprobs.size(); ++pi) {
        const Prob &p = dprobs[pi];
        const Side &sa = sides[jobs[p.job0].sa];
        const int ngrp = jobs[p.job0].ngrp;
        int64_t cost = 0;
        if (int8)
            cost = (int64_t)slices * (sa.Kp / 32 + 4);
        else
            for (int gi = 0; gi < ngrp; ++gi) cost += (int64_t)groups[jobs[p.job0].grp0 + gi].nseg * (sa.Kp / slice::BK) + 4;
        cost *= p.njobs;
        const int tmn = (int)((p.M + BTile::BM - 1) / BTile::BM), tnn = (int)((p.N + BTile::BN - 1) / BTile::BN);
        for (int tm = 0; tm < tmn; ++tm)
            for (int tn = 0; tn < tnn; ++tn) items.push_back(Item{(int)pi, tm, tn, (int)std::min<int64_t>(cost, INT_MAX)});
    }
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        split_batch_destroy(b);
        return bad(F4J_ERR_HIP, "device query failed");
    }
    int occ = 0;
    const hipError_t oe =
        int8 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bcrt_kernel, 256, 0)
        : carrier == SPLIT_FP16 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bgemm_kernel<_Float16, 64>, 256, 0)
                                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bgemm_kernel<__bf16, 256>, 256, 0);
    if (oe != hipSuccess || occ < 1) occ = 1;
    b->workers = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)prop.multiProcessorCount * occ,
                                                             (int64_t)items.size()));
    // Tiles dealt per XCD, longest first (xcd_schedule.h).
    std::vector<std::vector<Item>> per(b->workers);
    {
        std::vector<double> cost(items.size());
        std::vector<int> group(items.size());
        for (size_t i = 0; i < items.size(); ++i) {
            cost[i] = (double)items[i].cost;
            group[i] = items[i].prob;
        }
        const std::vector<std::vector<int>> idx = xcd_schedule(cost, group, b->workers);
        for (int w = 0; w < b->workers; ++w)
            for (int i : idx[w]) per[w].push_back(items[i]);
    }
    std::vector<Item> sched;
    std::vector<int> begin(b->workers + 1, 0);
    for (int wk = 0; wk < b->workers; ++wk) {
        begin[wk] = (int)sched.size();
        sched.insert(sched.end(), per[wk].begin(), per[wk].end());
    }
    begin[b->workers] = (int)sched.size();

    b->n_exp = (int64_t)exp_work.size();
    b->n_split = (int64_t)split_work.size();
    b->n_items = (int64_t)sched.size();
    b->n_res = (int64_t)res_work.size();
    if (b->n_exp > INT_MAX || b->n_split > INT_MAX || b->n_res > INT_MAX) {
        split_batch_destroy(b);
        return bad(F4J_ERR_UNSUPPORTED, "batch too large for one plan");
    }

    hipError_t e = hipSuccess;
    size_t &by = b->bytes;
    if (e == hipSuccess) e = upload(&b->d_sides, sides, &by);
    if (e == hipSuccess) e = upload(&b->d_jobs, jobs, &by);
    if (e == hipSuccess) e = upload(&b->d_probs, dprobs, &by);
    if (e == hipSuccess) e = upload(&b->d_groups, groups, &by);
    if (e == hipSuccess) e = upload(&b->d_items, sched, &by);
    if (e == hipSuccess) e = upload(&b->d_begin, begin, &by);
    if (e == hipSuccess) e = upload(&b->d_exp_work, exp_work, &by);
    if (e == hipSuccess) e = upload(&b->d_split_work, split_work, &by);
    if (e == hipSuccess) e = upload(&b->d_res_work, res_work, &by);
    if (e == hipSuccess && dig > 0) {
        const size_t db = (size_t)dig * (int8 ? 1 : 2);     // residue bytes, or 2-byte digits
        e = hipMalloc(&b->d_digits, db);
        by += db;
    }
    if (e == hipSuccess && eoff > 0) {
        e = hipMalloc((void **)&b->d_exps, (size_t)eoff * sizeof(int));
        by += (size_t)eoff * sizeof(int);
    }
    if (e == hipSuccess && !dprobs.empty()) {
        e = hipMalloc((void **)&b->d_err, dprobs.size() * sizeof(unsigned long long));
        by += dprobs.size() * sizeof(unsigned long long);
    }
    if (e == hipSuccess && b->d_err) e = hipMemset(b->d_err, 0, dprobs.size() * sizeof(unsigned long long));
    if (e != hipSuccess) {
        split_batch_destroy(b);
        return bad((e == hipErrorOutOfMemory) ? F4J_ERR_OUT_OF_MEMORY : F4J_ERR_HIP,
                   std::string("batched split plan: ") + hipGetErrorString(e));
    }
    return b;
}

hipError_t split_batch_enqueue(const SplitBatch *b, hipStream_t st, hipEvent_t ev_begin, hipEvent_t ev_end) {
    for (const BatchProblem &q : b->zero_k) {
        const int w = b->cplx ? 2 : 1;
        const hipError_t e = launch_zero_fill(q.C, q.M, q.N * w, q.ldc * w, st);
        if (e != hipSuccess) return e;
    }
    if (b->n_items == 0) return hipGetLastError();
    const bool fp16 = b->carrier == SPLIT_FP16;
    const int64_t nprob = (int64_t)std::count_if(b->prob_index.begin(), b->prob_index.end(), [](int v) { return v >= 0; });
    hipLaunchKernelGGL(fill_u64_kernel, dim3(64), dim3(256), 0, st, b->d_err, nprob);
    hipLaunchKernelGGL(bexp_kernel, dim3((unsigned)b->n_exp), dim3(256), 0, st, b->d_sides, b->d_exp_work, b->d_exps,
                       b->d_err);
    if (b->carrier == SPLIT_INT8) {
        const dim3 grid((unsigned)((b->n_split + 3) / 4));
        if (b->wide)
            hipLaunchKernelGGL(bcrt_split_kernel<true>, grid, dim3(256), 0, st, b->d_sides, b->d_split_work,
                               b->n_split, b->d_exps, (signed char *)b->d_digits, b->tb);
        else
            hipLaunchKernelGGL(bcrt_split_kernel<false>, grid, dim3(256), 0, st, b->d_sides, b->d_split_work,
                               b->n_split, b->d_exps, (signed char *)b->d_digits, b->tb);
        if (ev_begin) (void)hipEventRecord(ev_begin, st);
        hipLaunchKernelGGL(bcrt_kernel, dim3((unsigned)b->workers), dim3(256), 0, st, b->d_items, b->d_begin,
                           b->d_probs, b->d_jobs, b->d_sides, (const signed char *)b->d_digits, b->d_exps, b->tb,
                           (int)b->cplx);
        if (ev_end) (void)hipEventRecord(ev_end, st);
        if (b->n_res > 0)
            hipLaunchKernelGGL(bres_kernel, dim3((unsigned)b->n_res), dim3(256), 0, st, b->d_probs, b->d_res_work,
                               (int)b->cplx, b->opA, b->opB, b->d_err);
        return hipGetLastError();
    }
    if (fp16)
        launch_bsplit<10, _Float16>(b->S, dim3((unsigned)b->n_split), st, b->d_sides, b->d_split_work, b->d_exps,
                                    (_Float16 *)b->d_digits);
    else
        launch_bsplit<9, __bf16>(b->S, dim3((unsigned)b->n_split), st, b->d_sides, b->d_split_work, b->d_exps,
                                 (__bf16 *)b->d_digits);
    if (ev_begin) (void)hipEventRecord(ev_begin, st);
    if (fp16)
        hipLaunchKernelGGL((bgemm_kernel<_Float16, 64>), dim3((unsigned)b->workers), dim3(256), 0, st, b->d_items,
                           b->d_begin, b->d_probs, b->d_jobs, b->d_sides, b->d_groups, (const _Float16 *)b->d_digits,
                           b->d_exps, (int)b->cplx);
    else
        hipLaunchKernelGGL((bgemm_kernel<__bf16, 256>), dim3((unsigned)b->workers), dim3(256), 0, st, b->d_items,
                           b->d_begin, b->d_probs, b->d_jobs, b->d_sides, b->d_groups, (const __bf16 *)b->d_digits,
                           b->d_exps, (int)b->cplx);
    if (ev_end) (void)hipEventRecord(ev_end, st);
    if (b->n_res > 0) {
        hipLaunchKernelGGL(bres_kernel, dim3((unsigned)b->n_res), dim3(256), 0, st, b->d_probs, b->d_res_work,
                           (int)b->cplx, b->opA, b->opB, b->d_err);
    }
    return hipGetLastError();
}

hipError_t split_batch_errors(const SplitBatch *b, double *out, int64_t count, hipStream_t st) {
    const int64_t np = (int64_t)std::count_if(b->prob_index.begin(), b->prob_index.end(), [](int v) { return v >= 0; });
    std::vector<unsigned long long> bits(np, 0);
    if (np > 0) {
        hipError_t e = hipMemcpyAsync(bits.data(), b->d_err, np * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
    }
    for (int64_t i = 0; i < std::min<int64_t>(count, (int64_t)b->prob_index.size()); ++i) {
        const int k = b->prob_index[i];
        if (k < 0) out[i] = -1.0;
        else if (bits[k] >= NONFINITE_BITS) out[i] = std::numeric_limits<double>::quiet_NaN();   // NaN / Inf operands
        else out[i] = b->check_rows <= 0 ? -1.0 : __builtin_bit_cast(double, bits[k]);
    }
    return hipSuccess;
}

const unsigned long long *split_batch_err_bits(const SplitBatch *b) { return b->d_err; }
const std::vector<int> &split_batch_slots(const SplitBatch *b) { return b->prob_index; }
int split_batch_real_products(const SplitBatch *b) { return b->real_products; }
int split_batch_checked_rows(const SplitBatch *b) { return b->check_rows; }
size_t split_batch_bytes(const SplitBatch *b) { return b->bytes; }

void split_batch_destroy(SplitBatch *b) {
    if (!b) return;
    for (void *p : {(void *)b->d_sides, (void *)b->d_jobs, (void *)b->d_probs, (void *)b->d_groups, (void *)b->d_items,
                    (void *)b->d_begin, (void *)b->d_exp_work, (void *)b->d_split_work, (void *)b->d_res_work,
                    b->d_digits, (void *)b->d_exps, (void *)b->d_err})
        if (p) (void)hipFree(p);
    delete b;
}

}  // namespace f4j
