// Modular (CRT) split-precision FP64 GEMM on the INT8 matrix cores.
//
// No reference counterpart: Flag4j is FP64-only.  This is an alternative
// arithmetic for Matrix.mult (M:arrays/dense/Matrix.java:638-640,
// M: = src/main/java/org/flag4j/) selected by precision F4J_PREC_SPLIT_INT8.
// The decomposition follows the "Ozaki scheme II" idea (Ozaki, Uchino and
// Imamura, 2025): FP64 products from integer products modulo small coprime
// moduli, recombined by the Chinese remainder theorem.
//
// 1. Scaling.  Row i of A is scaled to t-bit fixed point relative to its
//    largest entry: A'_ik = rint(a_ik 2^(t - eA_i)), |A'| <= 2^t (columns of B
//    alike).  This rounding is the only approximation; t grows by about 4 bits
//    per extra modulus and may exceed 53, which helps rows whose entries span a
//    wide range (the small ones keep t - (eA_i - e(a_ik)) bits).
// 2. Residues.  For coprime moduli m_l <= 256, A'_ik mod m_l is stored as a
//    signed 8-bit residue (in [-128, 127]).
// 3. Products.  C_l = A_l B_l^T on v_mfma_i32_32x32x32_i8: |C_l| <= K 2^14,
//    exact in int32 for K < 2^17, then reduced to C_l mod m_l (one byte).
// 4. Reconstruction.  X = sum_l r_l W_l = A'B' (mod M), M = prod m_l; the
//    moduli are chosen with K 2^(2t) < M/4, so the representative of X in
//    (-M/4, M/4) is A'B' exactly.  W_l and M are cut into 40-bit chunks on a
//    common grid, so every partial sum S_p = sum_l r_l w_lp is an exact FP64
//    integer; q = rint(sum_p S_p scale_p / M) and D_p = S_p - q mu_p are exact;
//    C' = sum_p D_p scale_p is formed with the lower chunks in double-double,
//    so C' is rounded once, relative to itself.
//    C = C' 2^(eA_i + eB_j - 2t).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "amdgcn.h"
#include "crt_gemm.h"
#include "crt_layout.h"
#include "crt_residue.h"
#include "f4j_internal.h"
#include "gemm_tile.h"
#include "split_gemm.h"

namespace f4j {
namespace {

// Pairwise coprime moduli, largest first (256 = 2^8; 255 = 3*5*17; 253 = 11*23;
// 247 = 13*19; 217 = 7*31; the rest are primes).
constexpr int MODULI[CRT_MAX_MODULI] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                        223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

// ---- small unsigned bignum (base 2^32, little endian) for the tables ----
using Big = std::vector<uint32_t>;

Big big_from(uint64_t v) {
    Big b;
    while (v) {
        b.push_back((uint32_t)v);
        v >>= 32;
    }
    return b;
}
void big_trim(Big &b) {
    while (!b.empty() && b.back() == 0) b.pop_back();
}
Big big_mul_small(const Big &a, uint32_t x) {
    Big r(a.size() + 1, 0);
    uint64_t carry = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        const uint64_t v = (uint64_t)a[i] * x + carry;
        r[i] = (uint32_t)v;
        carry = v >> 32;
    }
    r[a.size()] = (uint32_t)carry;
    big_trim(r);
    return r;
}
Big big_div_small(const Big &a, uint32_t d, uint32_t *rem) {
    Big q(a.size(), 0);
    uint64_t r = 0;
    for (size_t i = a.size(); i-- > 0;) {
        const uint64_t cur = (r << 32) | a[i];
        q[i] = (uint32_t)(cur / d);
        r = cur % d;
    }
    big_trim(q);
    if (rem) *rem = (uint32_t)r;
    return q;
}
int big_bits(const Big &a) {
    if (a.empty()) return 0;
    int b = 32 * (int)(a.size() - 1);
    uint32_t top = a.back();
    while (top) {
        ++b;
        top >>= 1;
    }
    return b;
}
// bits [lo, lo + 40) of a (lo may be negative: bits below 0 are zero)
uint64_t big_bits40(const Big &a, int lo) {
    uint64_t v = 0;
    for (int k = 39; k >= 0; --k) {
        const int bit = lo + k;
        int b = 0;
        if (bit >= 0 && bit / 32 < (int)a.size()) b = (a[bit / 32] >> (bit % 32)) & 1;
        v = (v << 1) | (uint64_t)b;
    }
    return v;
}
double big_to_double(const Big &a) {
    const int nb = big_bits(a);
    if (nb <= 53) {
        double v = 0;
        for (size_t i = a.size(); i-- > 0;) v = v * 4294967296.0 + a[i];
        return v;
    }
    // top 53 bits, truncated, then rounded by the conversion of the next bits
    uint64_t top = 0;
    for (int k = nb - 1; k >= nb - 64; --k) top = (top << 1) | ((a[k / 32] >> (k % 32)) & 1);
    return std::ldexp((double)top, nb - 64);
}
int inv_mod(int a, int m) {
    int g = m, x = 0, x1 = 1, aa = a % m;
    while (aa) {
        const int q = g / aa;
        int t = g - q * aa;
        g = aa;
        aa = t;
        t = x - q * x1;
        x = x1;
        x1 = t;
    }
    return ((x % m) + m) % m;
}

int t_for(int n, int64_t K) {
    double lg = 0;
    for (int l = 0; l < n; ++l) lg += std::log2((double)MODULI[l]);
    const double lk = std::log2((double)std::max<int64_t>(K, 1));
    return (int)std::floor((lg - lk - 2.0) / 2.0 - 1e-9);
}

// ------------------------------------------------------------------------
// Residue slices: A'(x, k) mod m_l as signed bytes, fragment-major (below).
// ------------------------------------------------------------------------

using crt::pack4;


__device__ inline int exp_of(double v) {
    if (v == 0.0) return INT_MIN;
    int e;
    (void)frexp(v, &e);
    return e;
}

__global__ __launch_bounds__(256) void crt_fill_kernel(int *p, int64_t n, int v) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

// One launch before a call's exponent passes: the NaN / Inf flag to 0 (unless
// null) and, for
// operands read as strided lines (exponents merged with atomicMax), every job's
// exponent vector to INT_MIN (vector j of operand X at eX + j estride; nX = 0:
// nothing to fill).
__global__ __launch_bounds__(256) void crt_prep_kernel(int *flag, int *eA, int64_t nA, int *eB, int64_t nB,
                                                       int64_t estride) {
    if (flag && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *flag = 0;
    const int64_t j = blockIdx.y;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nA; i += (int64_t)gridDim.x * 256)
        eA[j * estride + i] = INT_MIN;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nB; i += (int64_t)gridDim.x * 256)
        eB[j * estride + i] = INT_MIN;
}

// Views of one operand's storage for the J jobs of a call: view j of element
// (re, im) is ca[j] re + cb[j] im (real data: the value itself, J = 1).
struct CrtCombos {
    double ca[CRT_MAX_JOBS], cb[CRT_MAX_JOBS];
};
__device__ inline double combo_value(const CrtOperand &o, const CrtCombos &cc, int j, double re, double im) {
    return o.cplx ? fma(cc.ca[j], re, cc.cb[j] * im) : re;
}
// The stored element (re, im) of line x, index k (im = 0 for real data).
__device__ inline void opnd_pair(const CrtOperand &o, int64_t x, int64_t k, double &re, double &im) {
    const double *p = o.X + x * o.sx + k * o.sk;
    re = p[0];
    im = o.cplx ? p[1] : 0.0;
}

// Line exponents (smallest E with |value| < 2^E over the line; INT_MIN for an
// all-zero line) of the J views, e[j estride + x].  Contiguous lines: one block
// per line, lanes along k.  A NaN or infinite entry sets *nonfinite (the
// residues cannot represent it).  Eight loads per lane are issued before any is
// used (one at a time kept too few bytes in flight: 3.9 TB/s).
constexpr int EXP_UNROLL = 8;
template <int J>
__global__ __launch_bounds__(256) void crt_exp_lines_kernel(CrtOperand o, CrtCombos cc, int64_t K, int *e,
                                                            int64_t estride, int *nonfinite) {
    const int64_t x = blockIdx.x;
    int m[J];
#pragma unroll
    for (int j = 0; j < J; ++j) m[j] = INT_MIN;
    bool bad = false;
    auto take = [&](double re, double im) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const double v = combo_value(o, cc, j, re, im);
            bad |= !__builtin_isfinite(v);
            m[j] = max(m[j], exp_of(v));
        }
    };
    int64_t k = threadIdx.x;
    for (; k + 256 * (EXP_UNROLL - 1) < K; k += 256 * EXP_UNROLL) {
        double re[EXP_UNROLL], im[EXP_UNROLL];
#pragma unroll
        for (int u = 0; u < EXP_UNROLL; ++u) opnd_pair(o, x, k + 256 * u, re[u], im[u]);
#pragma unroll
        for (int u = 0; u < EXP_UNROLL; ++u) take(re[u], im[u]);
    }
    for (; k < K; k += 256) {
        double re, im;
        opnd_pair(o, x, k, re, im);
        take(re, im);
    }
    if (bad) *nonfinite = 1;
    __shared__ int red[J][4];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        int mj = m[j];
        for (int off = 32; off > 0; off >>= 1) mj = max(mj, __shfl_xor(mj, off));
        if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = mj;
    }
    __syncthreads();
    if (threadIdx.x < J)
        e[threadIdx.x * estride + x] =
            max(max(red[threadIdx.x][0], red[threadIdx.x][1]), max(red[threadIdx.x][2], red[threadIdx.x][3]));
}

// Strided lines: thread per line (lanes along the lines), one k-chunk per
// block row, merged with atomicMax into e (filled with INT_MIN first).
template <int J>
__global__ __launch_bounds__(256) void crt_exp_cols_kernel(CrtOperand o, CrtCombos cc, int64_t nx, int64_t K,
                                                           int64_t kchunk, int *e, int64_t estride, int *nonfinite) {
    const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (x >= nx) return;
    const int64_t k0 = (int64_t)blockIdx.y * kchunk, k1 = min(K, k0 + kchunk);
    int m[J];
#pragma unroll
    for (int j = 0; j < J; ++j) m[j] = INT_MIN;
    bool bad = false;
    auto take = [&](double re, double im) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const double v = combo_value(o, cc, j, re, im);
            bad |= !__builtin_isfinite(v);
            m[j] = max(m[j], exp_of(v));
        }
    };
    int64_t k = k0;
    for (; k + EXP_UNROLL <= k1; k += EXP_UNROLL) {
        double re[EXP_UNROLL], im[EXP_UNROLL];
#pragma unroll
        for (int u = 0; u < EXP_UNROLL; ++u) opnd_pair(o, x, k + u, re[u], im[u]);
#pragma unroll
        for (int u = 0; u < EXP_UNROLL; ++u) take(re[u], im[u]);
    }
    for (; k < k1; ++k) {
        double re, im;
        opnd_pair(o, x, k, re, im);
        take(re, im);
    }
    if (bad) *nonfinite = 1;
#pragma unroll
    for (int j = 0; j < J; ++j)
        if (m[j] != INT_MIN) atomicMax(e + j * estride + x, m[j]);
}

// Contiguous lines (a line is a row of the stored matrix): one block is 32
// lines x 128 k, staged through LDS with coalesced reads along k; thread t then
// takes line t % 32, 16 k at 16 ((t / 32) % 2) + 32 (t / 64) and writes one
// 16-byte fragment piece per modulus (a wave writes 1 KB).
template <bool WIDE, int J>
__global__ __launch_bounds__(256) void crt_split_kernel(CrtOperand o, CrtCombos cc, int64_t nx, int64_t K, const int *e,
                                                        int64_t estride, signed char *out, int64_t ostride, int64_t Xp,
                                                        int64_t Kp, CrtTables tb) {
    __shared__ double tile[32][129];
    const int64_t k0 = (int64_t)blockIdx.x * 128, x0 = (int64_t)blockIdx.y * 32;
    const int t = threadIdx.x;
    for (int j = 0; j < J; ++j) {      // view j of the operand (one job)
        {
            const int c = t & 127;
            for (int r = t >> 7; r < 32; r += 2) {
                const int64_t x = x0 + r, k = k0 + c;
                double re = 0.0, im = 0.0;
                if (x < nx && k < K) opnd_pair(o, x, k, re, im);
                tile[r][c] = combo_value(o, cc, j, re, im);
            }
        }
        __syncthreads();
        const int r = crt::split_line(t), kb = crt::split_k(t);
        const int64_t x = x0 + r;
        const int ex = x < nx ? e[j * estride + x] : INT_MIN;
        // scaled integers, split once (v = vh 2^26 + vl) for the per-modulus residues
        const int sh = crt::scaled_shift(ex, tb.t);
        double vh[16], vl[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const double v = crt::scaled_int_sh(tile[r][kb + u], sh);
            if (WIDE) {
                vh[u] = v;               // kept whole: the wide step reduces it per modulus
                vl[u] = 0.0;
            } else {
                crt::residue_split(v, vh[u], vl[u]);
                vl[u] = crt::residue_bias(vl[u]);     // the biased form (crt_residue.h)
            }
        }
        const int64_t KC = Kp >> 5;
        int4 *dst = (int4 *)(out + j * ostride + crt::frag_offset(x, k0 + kb, KC));
        const int64_t lstride = Xp * Kp / 16;
        for (int l = 0; l < tb.n; ++l) {
            const double m = (double)tb.m[l], mi = tb.minv[l], c26 = tb.c26[l], cq = crt::residue_cq(mi);
            unsigned rs[16];
#pragma unroll
            for (int u = 0; u < 16; ++u)
                rs[u] = WIDE ? crt::residue_wide_biased(vh[u], m, mi, cq, c26)
                             : crt::residue_biased(vh[u], vl[u], m, mi, cq, c26);
            int4 w;
            w.x = (int)crt::pack4_biased(rs[0], rs[1], rs[2], rs[3]);
            w.y = (int)crt::pack4_biased(rs[4], rs[5], rs[6], rs[7]);
            w.z = (int)crt::pack4_biased(rs[8], rs[9], rs[10], rs[11]);
            w.w = (int)crt::pack4_biased(rs[12], rs[13], rs[14], rs[15]);
            dst[l * lstride] = w;
        }
        if (j + 1 < J) __syncthreads();    // the tile is reloaded for the next view
    }
}

// Strided lines (a line is a column of the stored matrix): one block is 128
// lines x 32 k, so every k row of the block is read as 1 KB of contiguous
// memory; thread t then takes line t % 128 and 16 k at 16 (t / 128).
template <bool WIDE, int J>
__global__ __launch_bounds__(256) void crt_split_cols_kernel(CrtOperand o, CrtCombos cc, int64_t nx, int64_t K,
                                                             const int *e, int64_t estride, signed char *out,
                                                             int64_t ostride, int64_t Xp, int64_t Kp, CrtTables tb) {
    __shared__ double tile[128][33];
    const int64_t k0 = (int64_t)blockIdx.x * 32, x0 = (int64_t)blockIdx.y * 128;
    const int t = threadIdx.x;
    for (int j = 0; j < J; ++j) {      // view j of the operand (one job)
        {
            const int c = t & 127;
            for (int r = t >> 7; r < 32; r += 2) {
                const int64_t x = x0 + c, k = k0 + r;
                double re = 0.0, im = 0.0;
                if (x < nx && k < K) opnd_pair(o, x, k, re, im);
                tile[c][r] = combo_value(o, cc, j, re, im);
            }
        }
        __syncthreads();
        const int xl = crt::split_cols_line(t), kb = crt::split_cols_k(t);
        const int64_t x = x0 + xl;
        const int ex = x < nx ? e[j * estride + x] : INT_MIN;
        const int sh = crt::scaled_shift(ex, tb.t);
        double vh[16], vl[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const double v = crt::scaled_int_sh(tile[xl][kb + u], sh);
            if (WIDE) {
                vh[u] = v;
                vl[u] = 0.0;
            } else {
                crt::residue_split(v, vh[u], vl[u]);
                vl[u] = crt::residue_bias(vl[u]);     // the biased form (crt_residue.h)
            }
        }
        const int64_t KC = Kp >> 5;
        int4 *dst = (int4 *)(out + j * ostride + crt::frag_offset(x, k0 + kb, KC));
        const int64_t lstride = Xp * Kp / 16;
        for (int l = 0; l < tb.n; ++l) {
            const double m = (double)tb.m[l], mi = tb.minv[l], c26 = tb.c26[l], cq = crt::residue_cq(mi);
            unsigned rs[16];
#pragma unroll
            for (int u = 0; u < 16; ++u)
                rs[u] = WIDE ? crt::residue_wide_biased(vh[u], m, mi, cq, c26)
                             : crt::residue_biased(vh[u], vl[u], m, mi, cq, c26);
            int4 w;
            w.x = (int)crt::pack4_biased(rs[0], rs[1], rs[2], rs[3]);
            w.y = (int)crt::pack4_biased(rs[4], rs[5], rs[6], rs[7]);
            w.z = (int)crt::pack4_biased(rs[8], rs[9], rs[10], rs[11]);
            w.w = (int)crt::pack4_biased(rs[12], rs[13], rs[14], rs[15]);
            dst[l * lstride] = w;
        }
        if (j + 1 < J) __syncthreads();    // the tile is reloaded for the next view
    }
}

// ------------------------------------------------------------------------
// Residue products: R[l][i][j] = (A_l B_l^T)_ij mod m_l, one byte.
// v_mfma_i32_32x32x32_i8: A/B lane l holds row l%32, k = 16(l/32)+v; result
// item r of lane l is row 8(r/4) + 4(l/32) + r%4, column l%32.
// ------------------------------------------------------------------------


constexpr int CBM = crt::BM, CBN = crt::BN, CBK = crt::BK;

struct CrtGemmArgs {
    const signed char *A, *B;                     // fragment-major residues, n slices of Mp x Kp and Np x Kp
    int64_t Mp, Np, Kp, M, N;
    unsigned char *R;                             // [n][plane], each plane M x N
    int64_t plane;
    int64_t job_stride;                           // bytes between the jobs' workspaces (blockIdx.y = job n + l)
};

// Block tile BM x 256 (BM = 128 RB), eight waves of 32 RB x 128 (RB x 4 MFMA
// blocks, 64 RB int32 accumulators per lane): wave w takes row group w % 4 and
// column half w / 4.
// Both operands reach LDS by buffer loads that write LDS directly, each
// fragment once per block (waves 0-3 load A's, 4-7 B's), through a ring of
// STAGES steps of SK k bytes, loaded up to STAGES steps ahead of the MFMAs,
// one barrier per step.  Loading A once per block rather than once per wave into registers
// takes a third of the bytes off the CU's vector-memory return path, which
// was the busiest unit; gfx950's LDS carries the extra fragment reads.  With
// the fragment-major layout every load and every fragment read moves 1 KB of
// contiguous memory.
__global__ __launch_bounds__(512, 2) void crt_gemm_kernel(CrtGemmArgs a, CrtTables tb) {
    __shared__ __attribute__((aligned(1024))) signed char lds[crt::STAGES * crt::STAGE];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform: keeps descriptors scalar
    const int l = blockIdx.y % tb.n;
    const int64_t jo = (int64_t)(blockIdx.y / tb.n) * a.job_stride;
    int64_t tm, tn;
    tile_of(blockIdx.x, a.Mp / CBM, a.Np / CBN, tm, tn);
    const int64_t m0 = (int64_t)__builtin_amdgcn_readfirstlane((int)tm) * CBM;
    const int64_t n0 = (int64_t)__builtin_amdgcn_readfirstlane((int)tn) * CBN;

    i32x16 acc[crt::RB][4];
#pragma unroll
    for (int i = 0; i < crt::RB; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0;

    const int KC = (int)(a.Kp >> 5);
    // this wave's operand: A's slice (waves 0-3) or B's (waves 4-7) of modulus l
    const bool is_a = crt::loads_a(wave);
    const signed char *src = is_a ? a.A + jo + (int64_t)l * a.Mp * a.Kp : a.B + jo + (int64_t)l * a.Np * a.Kp;
    const int64_t x0 = is_a ? m0 : n0;
    const int nt = (int)(a.Kp / crt::SK);

    auto issue = [&](int slot, int t) {          // step t into stage slot % STAGES
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(src + crt::load_base(x0, wave, t, KC)), (short)0, 0x7FFFFFF0, 0x00020000);
        auto loads = [&](auto count) {
#pragma unroll
            for (int p = 0; p < decltype(count)::value; ++p)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void *)(lds + crt::load_lds(slot, wave, p)), 16,
                                                         crt::load_voff(lane), crt::load_soff(p, KC), 0, 0);
        };
        if (is_a)                                // uniform: A waves load LOADS_A fragments, B waves LOADS_B
            loads(std::integral_constant<int, crt::LOADS_A>{});
        else
            loads(std::integral_constant<int, crt::LOADS_B>{});
    };
    // Pipelined across steps: the MFMAs of a step's last k step run while the
    // waves meet at the barrier, issue the loads STAGES steps ahead and read the
    // next step's first fragments, so the next step's first MFMAs find their
    // operands ready.  At the barrier of step t every wave has read all of stage
    // t % STAGES and has its own loads of step t + 1 retired, so after it step
    // t + 1 may be read and step t + STAGES may overwrite stage t % STAGES.
    // Every step issues the same loads (past the last step they repeat step
    // nt - 1 into the free stage); with two stages the wave's loads of step
    // t + 1 are its most recent ones, so vmcnt(0) retires them.
    static_assert(crt::KB % 2 == 0, "even k steps per stage");
    static_assert(crt::STAGES == 2, "the waits below assume two stages");
    i32x4 fa[2][crt::RB], fb[2][4];
#pragma unroll
    for (int t = 0; t < crt::STAGES - 1; ++t) issue(t, min(t, nt - 1));
    F4J_WAIT_VMCNT(0);
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < crt::RB; ++i) fa[0][i] = *(const i32x4 *)(lds + crt::a_frag(0, wave, i, 0, lane));
#pragma unroll
    for (int j = 0; j < 4; ++j) fb[0][j] = *(const i32x4 *)(lds + crt::b_frag(0, wave, j, 0, lane));
    __builtin_amdgcn_sched_barrier(0);
    issue(crt::STAGES - 1, min(crt::STAGES - 1, nt - 1));
    auto mfmas = [&](int buf) {
#pragma unroll
        for (int i = 0; i < crt::RB; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[buf][i], fb[buf][j], acc[i][j], 0, 0, 0);
    };
    for (int t = 0; t < nt; ++t) {
#pragma unroll
        for (int ks = 0; ks + 1 < crt::KB; ++ks) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < crt::RB; ++i)
                fa[(ks + 1) & 1][i] = *(const i32x4 *)(lds + crt::a_frag(t, wave, i, ks + 1, lane));
#pragma unroll
            for (int j = 0; j < 4; ++j)
                fb[(ks + 1) & 1][j] = *(const i32x4 *)(lds + crt::b_frag(t, wave, j, ks + 1, lane));
            mfmas(ks & 1);
            // one fragment read after each of the first MFMAs: the reads of k step ks + 1
            // then have the rest of the k step's MFMAs to arrive (2-3% over letting
            // the compiler gather them after the MFMAs)
#pragma unroll
            for (int q = 0; q < crt::RB + 4; ++q) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);     // MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);     // LDS read
            }
            __builtin_amdgcn_sched_group_barrier(0x008, 4 * crt::RB - (crt::RB + 4), 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        F4J_WAIT_VMCNT(0);                        // this wave's loads of step t + 1 are in LDS
        __builtin_amdgcn_s_waitcnt(0xC07F);      // lgkmcnt(0): its reads of stage t % STAGES are done
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(t + crt::STAGES, min(t + crt::STAGES, nt - 1));   // into stage t % STAGES
        // step t + 1's first fragments (past the last step: the repeated stage, never used),
        // then step t's last k step; reads and MFMAs alternate, the loads go after the first MFMA
#pragma unroll
        for (int i = 0; i < crt::RB; ++i) fa[0][i] = *(const i32x4 *)(lds + crt::a_frag(t + 1, wave, i, 0, lane));
#pragma unroll
        for (int j = 0; j < 4; ++j) fb[0][j] = *(const i32x4 *)(lds + crt::b_frag(t + 1, wave, j, 0, lane));
        mfmas((crt::KB - 1) & 1);
#pragma unroll
        for (int q = 0; q < crt::RB + 4; ++q) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 1);         // MFMA
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 1);         // LDS read
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 1);
        __builtin_amdgcn_sched_group_barrier(0x020, crt::LOADS_A, 1);  // global loads into LDS
        __builtin_amdgcn_sched_group_barrier(0x008, 4 * crt::RB - (crt::RB + 4) - 1, 1);
        __builtin_amdgcn_sched_barrier(0);
    }
    F4J_WAIT_VMCNT(0);                            // no LDS write in flight when the block ends

    // acc mod m (crt::product_residue: in [0, m], m kept for exact multiples), a
    // lane's 16 residues of a block as one 16-byte store into the padded plane
    // (crt::res_block): no per-entry range checks; the reconstruction skips padding.
    const double m = (double)tb.m[l];
    const double minv = tb.minv[l];
    unsigned char *R = a.R + jo + (int64_t)l * a.plane + 16 * lane;
    const int64_t NB = a.Np >> 5;
    const int64_t rw = m0 + crt::wave_row(wave), cw = n0 + crt::wave_col(wave);
#pragma unroll
    for (int i = 0; i < crt::RB; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int rs[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) rs[r] = crt::product_residue(acc[i][j][r], m, minv);
            int4 w;
            w.x = (int)pack4(rs[0], rs[1], rs[2], rs[3]);
            w.y = (int)pack4(rs[4], rs[5], rs[6], rs[7]);
            w.z = (int)pack4(rs[8], rs[9], rs[10], rs[11]);
            w.w = (int)pack4(rs[12], rs[13], rs[14], rs[15]);
            *(int4 *)(R + crt::res_block(rw + 32 * i, cw + 32 * j, NB)) = w;
        }
}

// ------------------------------------------------------------------------
// Reconstruction: C_ij = 2^(eA_i + eB_j - 2t) * CRT(R[.][i][j]).
// ------------------------------------------------------------------------

// Thread -> one lane's 16 entries of a 32 x 32 block of the product planes
// (crt::res_block; a wave is one block): one 16-byte load per modulus, all
// issued before any is used, then four rounds of four entries (rows
// crt::res_row, column crt::res_col), so only four entries' chunk sums are live.
// NW = residue words loaded (the moduli rounded up to a multiple of four:
// words past n re-read plane 0, a cache hit, unused), which bounds the
// registers by the moduli in use.
// Output: real C (cplx = 0): C_ij = value; complex C (interleaved):
// Re C_ij (+)= cre value, Im C_ij (+)= cim value (first: assign).
template <int NW>
__global__ __launch_bounds__(256) void crt_reconstruct_kernel(const unsigned char *R, int64_t plane, int64_t NB,
                                                              int64_t nblocks, int64_t M, int64_t N, const int *eA,
                                                              const int *eB, double *C, int64_t ldc, CrtTables tb,
                                                              int cplx, double cre, double cim, int first) {
    const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= nblocks) return;
    const int64_t row0 = (b / NB) * 32, col0 = (b % NB) * 32;
    if (row0 >= M || col0 >= N) return;           // padding block (uniform per wave)
    int4 word[NW];
#pragma unroll
    for (int l = 0; l < NW; ++l)
        word[l] = *(const int4 *)(R + (int64_t)(l < tb.n ? l : 0) * plane + (b << 10) + 16 * lane);
    const int64_t j = col0 + crt::res_col(lane);
    const bool col_ok = j < N;
    const int eb = eB[col_ok ? j : N - 1];
#pragma unroll
    for (int q = 0; q < 4; ++q) {                 // items 4q .. 4q + 3 (word component q)
        double S[CRT_MAX_CHUNKS][4];
#pragma unroll
        for (int p = 0; p < CRT_MAX_CHUNKS; ++p)
#pragma unroll
            for (int e = 0; e < 4; ++e) S[p][e] = 0.0;
#pragma unroll
        for (int l = 0; l < NW; ++l) {
            if (l < tb.n) {      // uniform; no early exit, so word[] keeps static indices (registers)
                const unsigned wq = (unsigned)(q == 0 ? word[l].x : q == 1 ? word[l].y : q == 2 ? word[l].z : word[l].w);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const double r = (double)((wq >> (8 * e)) & 255u);
#pragma unroll
                    for (int p = 0; p < CRT_MAX_CHUNKS; ++p) S[p][e] = fma(r, tb.w[l][p], S[p][e]);   // exact integers
                }
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t i = row0 + crt::res_row(4 * q + e, lane);
            if (!col_ok || i >= M) continue;
            const int ea = eA[i];
            double out = 0.0;
            if (ea != INT_MIN && eb != INT_MIN)
                out = ldexp(crt_recombine(S[0][e], S[1][e], S[2][e], S[3][e], tb), ea + eb - 2 * tb.t);
            if (cplx) {
                double *c = C + 2 * (i * ldc + j);
                c[0] = first ? cre * out : c[0] + cre * out;
                c[1] = first ? cim * out : c[1] + cim * out;
            } else {
                C[i * ldc + j] = out;
            }
        }
    }
}

// All J jobs of a complex product in one pass (jobs in consecutive workspaces
// `jstride` bytes apart, with their own residue planes and exponents): for each
// entry, job j's value v_j is reconstructed as above, and Re C = sum_j cre_j v_j,
// Im C = sum_j cim_j v_j are summed in job order from the first job's terms --
// the order in which successive one-job launches would apply them -- and C is
// written once.
struct CrtJobCoef {
    double cre[CRT_MAX_JOBS], cim[CRT_MAX_JOBS];
};
template <int NW, int J>
__global__ __launch_bounds__(256) void crt_reconstruct_jobs_kernel(const unsigned char *R, int64_t plane, int64_t NB,
                                                                   int64_t nblocks, int64_t M, int64_t N,
                                                                   const int *eA, const int *eB, int64_t jstride,
                                                                   double *C, int64_t ldc, CrtTables tb,
                                                                   CrtJobCoef jc) {
    const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= nblocks) return;
    const int64_t row0 = (b / NB) * 32, col0 = (b % NB) * 32;
    if (row0 >= M || col0 >= N) return;           // padding block (uniform per wave)
    const int64_t j = col0 + crt::res_col(lane);
    const bool col_ok = j < N;
    // the lane's 16 entries of C (result items r = 0 .. 15), summed over the jobs; each
    // job's residue words are loaded whole (one 16-byte load per modulus) before use
    // (rolling the jobs or the item rounds to save registers measured slower: 145 us
    // against 110 at 2048 x 1536 x 2048 with 3 jobs and 16 moduli)
    double re[16], im[16];
#pragma unroll
    for (int jb = 0; jb < J; ++jb) {
        const unsigned char *Rj = R + jb * jstride;
        const int *eAj = (const int *)((const char *)eA + jb * jstride);
        const int *eBj = (const int *)((const char *)eB + jb * jstride);
        int4 word[NW];
#pragma unroll
        for (int l = 0; l < NW; ++l)
            word[l] = *(const int4 *)(Rj + (int64_t)(l < tb.n ? l : 0) * plane + (b << 10) + 16 * lane);
        const int eb = eBj[col_ok ? j : N - 1];
#pragma unroll
        for (int q = 0; q < 4; ++q) {             // items 4q .. 4q + 3 (word component q)
            double S[CRT_MAX_CHUNKS][4];
#pragma unroll
            for (int p = 0; p < CRT_MAX_CHUNKS; ++p)
#pragma unroll
                for (int e = 0; e < 4; ++e) S[p][e] = 0.0;
#pragma unroll
            for (int l = 0; l < NW; ++l) {
                if (l < tb.n) {  // uniform; no early exit, so word[] keeps static indices (registers)
                    const unsigned wq =
                        (unsigned)(q == 0 ? word[l].x : q == 1 ? word[l].y : q == 2 ? word[l].z : word[l].w);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const double r = (double)((wq >> (8 * e)) & 255u);
#pragma unroll
                        for (int p = 0; p < CRT_MAX_CHUNKS; ++p) S[p][e] = fma(r, tb.w[l][p], S[p][e]);
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = row0 + crt::res_row(4 * q + e, lane);
                double out = 0.0;
                if (col_ok && i < M) {
                    const int ea = eAj[i];
                    if (ea != INT_MIN && eb != INT_MIN)
                        out = ldexp(crt_recombine(S[0][e], S[1][e], S[2][e], S[3][e], tb), ea + eb - 2 * tb.t);
                }
                const int it = 4 * q + e;
                re[it] = jb == 0 ? jc.cre[0] * out : re[it] + jc.cre[jb] * out;
                im[it] = jb == 0 ? jc.cim[0] * out : im[it] + jc.cim[jb] * out;
            }
        }
    }
    if (!col_ok) return;
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int64_t i = row0 + crt::res_row(it, lane);
        if (i < M) {
            double *c = C + 2 * (i * ldc + j);
            c[0] = re[it];
            c[1] = im[it];
        }
    }
}

}  // namespace

int crt_moduli_for_bits(int bits, int64_t K) {
    for (int n = 2; n <= CRT_MAX_MODULI; ++n)
        if (t_for(n, K) >= bits) return n;
    return CRT_MAX_MODULI;
}

int crt_default_moduli(int64_t K) { return crt_moduli_for_bits(53, K); }

int crt_bits(int n, int64_t K) { return t_for(n, K); }

bool crt_tables(int n, int64_t K, CrtTables *out) {
    if (n < 2 || n > CRT_MAX_MODULI) return false;
    const int t = t_for(n, K);
    if (t < 1) return false;
    CrtTables tb{};
    tb.n = n;
    tb.t = t;
    Big Mb = big_from(1);
    for (int l = 0; l < n; ++l) Mb = big_mul_small(Mb, (uint32_t)MODULI[l]);
    const int T = big_bits(Mb);
    tb.P = (T + 39) / 40;
    if (tb.P > CRT_MAX_CHUNKS) return false;
    const int s = 40 * tb.P - T;          // W' = W 2^s fills exactly P chunks of 40 bits
    const double Md = big_to_double(Mb);
    for (int p = 0; p < CRT_MAX_CHUNKS; ++p) {
        const int lo = 40 * (tb.P - 1 - p);
        if (p < tb.P) {
            tb.scale[p] = std::ldexp(1.0, lo - s);
            tb.mu[p] = (double)big_bits40(Mb, lo - s);
            tb.c[p] = tb.scale[p] / Md;
        } else {
            tb.scale[p] = 0.0;
            tb.mu[p] = 0.0;
            tb.c[p] = 0.0;
        }
    }
    for (int l = 0; l < CRT_MAX_MODULI; ++l) {
        for (int p = 0; p < CRT_MAX_CHUNKS; ++p) tb.w[l][p] = 0.0;
        tb.m[l] = MODULI[l];
        tb.minv[l] = 1.0 / MODULI[l];
        tb.c26[l] = (double)((1 << 26) % MODULI[l]);
        if (l >= n) continue;
        uint32_t rem = 0;
        const Big Ml = big_div_small(Mb, (uint32_t)MODULI[l], nullptr);
        (void)big_div_small(Ml, (uint32_t)MODULI[l], &rem);
        const Big W = big_mul_small(Ml, (uint32_t)inv_mod((int)rem, MODULI[l]));
        for (int p = 0; p < tb.P; ++p) tb.w[l][p] = (double)big_bits40(W, 40 * (tb.P - 1 - p) - s);
    }
    *out = tb;
    return true;
}

bool plan_crt(int64_t M, int64_t N, int64_t K, int n, int check_rows, CrtPlan *plan) {
    CrtPlan pl;
    if (!crt_tables(n, K, &pl.tb)) return false;
    pl.Mp = (M + CBM - 1) / CBM * CBM;
    pl.Np = (N + CBN - 1) / CBN * CBN;
    pl.Kp = (K + CBK - 1) / CBK * CBK;
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    pl.off_ra = 0;
    pl.off_rb = al((size_t)n * pl.Mp * pl.Kp);
    pl.plane = (int64_t)(pl.Mp * pl.Np);          // crt::res_block layout, a multiple of 2^16
    pl.off_res = pl.off_rb + al((size_t)n * pl.Np * pl.Kp);
    pl.off_ea = pl.off_res + (size_t)n * pl.plane;
    pl.off_eb = pl.off_ea + al((size_t)M * 4);
    pl.off_err = pl.off_eb + al((size_t)N * 4);
    pl.off_flag = pl.off_err + 128;
    pl.off_part = pl.off_err + 256;
    pl.bytes = pl.off_part + split_residual_bytes(N, K, (int)std::min<int64_t>(check_rows, M));
    *plan = pl;
    return true;
}

CrtOperand crt_operand(const double *X, int64_t ld, bool lines, bool cplx, double ca, double cb) {
    CrtOperand o{};
    const int w = cplx ? 2 : 1;
    o.X = X;
    o.sx = (lines ? ld : 1) * w;
    o.sk = (lines ? 1 : ld) * w;
    o.ca = ca;
    o.cb = cb;
    o.cplx = cplx;
    o.lines = lines;
    return o;
}

namespace {

// Enqueue J jobs on the storage of A and B (job j's operand views ccA/ccB, its
// workspace crt_job_stride(pl) bytes after job j - 1's): exponents and residues of
// every view from one kernel per operand and pass (each pass reads the operand
// once, its later views from L2), one product launch over all jobs' moduli, one
// reconstruction (J = 1: the one-job kernel with `out`; J > 1: complex C written
// once from the jobs' coefficients `jc`).
template <int J>
hipError_t enqueue_crt(const CrtPlan &pl, const CrtOperand &oa, const CrtOperand &ob, const CrtCombos &ccA,
                       const CrtCombos &ccB, int64_t M, int64_t N, int64_t K, double *C, int64_t ldc,
                       const CrtOutput &out, const CrtJobCoef &jc, void *ws, hipStream_t stream, hipEvent_t ev_begin,
                       hipEvent_t ev_end, const CrtSide *side) {
    const int64_t js = (int64_t)crt_job_stride(pl);
    char *w = (char *)ws;
    signed char *ra = (signed char *)(w + pl.off_ra), *rb = (signed char *)(w + pl.off_rb);
    unsigned char *res = (unsigned char *)(w + pl.off_res);
    int *ea = (int *)(w + pl.off_ea), *eb = (int *)(w + pl.off_eb);
    int *nonfinite = (int *)(w + pl.off_flag);
    auto exps = [&](const CrtOperand &o, const CrtCombos &cc, int64_t nx, int *e, hipStream_t stream) {
        if (nx <= 0) return;
        if (o.lines) {
            hipLaunchKernelGGL((crt_exp_lines_kernel<J>), dim3((unsigned)nx), dim3(256), 0, stream, o, cc, K, e,
                               js / 4, nonfinite);
        } else {
            // k chunks sized for at least ~2048 workgroups (a fixed 256 gave 48
            // workgroups for 2048 columns and K = 1536: 130 us instead of ~10)
            const int64_t col_blocks = (nx + 255) / 256;
            const int64_t kchunk =
                std::max<int64_t>(EXP_UNROLL, std::min<int64_t>(256, (K * col_blocks / 2048) / EXP_UNROLL * EXP_UNROLL));
            // e was filled with INT_MIN by crt_prep_kernel
            hipLaunchKernelGGL((crt_exp_cols_kernel<J>),
                               dim3((unsigned)((nx + 255) / 256), (unsigned)((K + kchunk - 1) / kchunk)), dim3(256), 0,
                               stream, o, cc, nx, K, kchunk, e, js / 4, nonfinite);
        }
    };
    // the flag (a product's first call only: later ones keep what earlier exponent
    // passes flagged) and the INT_MIN start of strided-line exponents, in one launch
    if (out.first || !oa.lines || !ob.lines) {
        const int64_t nA = oa.lines ? 0 : M, nB = ob.lines ? 0 : N;
        hipLaunchKernelGGL(crt_prep_kernel, dim3(64, J), dim3(256), 0, stream, out.first ? nonfinite : nullptr, ea,
                           nA, eb, nB, js / 4);
    }
    const bool wide = pl.tb.t > 58;
    auto split = [&](const CrtOperand &o, const CrtCombos &cc, int64_t nx, int64_t Xp, const int *ex,
                     signed char *dst, hipStream_t stream) {
        if (o.lines) {
            const dim3 grid((unsigned)(pl.Kp / 128), (unsigned)(Xp / 32));
            if (wide)
                hipLaunchKernelGGL((crt_split_kernel<true, J>), grid, dim3(256), 0, stream, o, cc, nx, K, ex, js / 4,
                                   dst, js, Xp, pl.Kp, pl.tb);
            else
                hipLaunchKernelGGL((crt_split_kernel<false, J>), grid, dim3(256), 0, stream, o, cc, nx, K, ex, js / 4,
                                   dst, js, Xp, pl.Kp, pl.tb);
        } else {
            const dim3 gc((unsigned)(pl.Kp / 32), (unsigned)(Xp / 128));
            if (wide)
                hipLaunchKernelGGL((crt_split_cols_kernel<true, J>), gc, dim3(256), 0, stream, o, cc, nx, K, ex,
                                   js / 4, dst, js, Xp, pl.Kp, pl.tb);
            else
                hipLaunchKernelGGL((crt_split_cols_kernel<false, J>), gc, dim3(256), 0, stream, o, cc, nx, K, ex,
                                   js / 4, dst, js, Xp, pl.Kp, pl.tb);
        }
    };
    hipStream_t sb = stream;
    if (side) {
        sb = side->stream;
        if (hipError_t e = hipEventRecord(side->fork, stream); e != hipSuccess) return e;
        if (hipError_t e = hipStreamWaitEvent(sb, side->fork, 0); e != hipSuccess) return e;
    }
    // the two operands' launches interleaved, so that both streams get work as soon as
    // the host has issued their first kernel (at mid sizes the issue time of a launch
    // is a good part of a kernel's time)
    exps(ob, ccB, N, eb, sb);
    exps(oa, ccA, M, ea, stream);
    split(ob, ccB, N, pl.Np, eb, rb, sb);
    if (side) {
        if (hipError_t e = hipEventRecord(side->join, sb); e != hipSuccess) return e;
    }
    split(oa, ccA, M, pl.Mp, ea, ra, stream);
    if (side) {
        if (hipError_t e = hipStreamWaitEvent(stream, side->join, 0); e != hipSuccess) return e;
    }
    CrtGemmArgs a{ra, rb, pl.Mp, pl.Np, pl.Kp, M, N, res, pl.plane, js};
    if (ev_begin) (void)hipEventRecord(ev_begin, stream);
    hipLaunchKernelGGL(crt_gemm_kernel, dim3((unsigned)((pl.Mp / CBM) * (pl.Np / CBN)), (unsigned)(pl.tb.n * J)),
                       dim3(64 * crt::WAVES), 0, stream, a, pl.tb);
    if (ev_end) (void)hipEventRecord(ev_end, stream);
    const int64_t NB = pl.Np / 32, nblocks = (pl.Mp / 32) * NB;
    const dim3 rgrid((unsigned)((nblocks + 3) / 4));
    if constexpr (J == 1) {
#define F4J_RECONSTRUCT(NW)                                                                                        \
    hipLaunchKernelGGL(crt_reconstruct_kernel<NW>, rgrid, dim3(256), 0, stream, res, pl.plane, NB, nblocks, M, N, ea, \
                       eb, C, ldc, pl.tb, (int)out.cplx, out.cre, out.cim, (int)out.first)
        if (pl.tb.n <= 8) F4J_RECONSTRUCT(8);
        else if (pl.tb.n <= 12) F4J_RECONSTRUCT(12);
        else if (pl.tb.n <= 16) F4J_RECONSTRUCT(16);
        else F4J_RECONSTRUCT(CRT_MAX_MODULI);
#undef F4J_RECONSTRUCT
    } else {
#define F4J_RECONSTRUCT(NW)                                                                                          \
    hipLaunchKernelGGL((crt_reconstruct_jobs_kernel<NW, J>), rgrid, dim3(256), 0, stream, res, pl.plane, NB, nblocks, M, \
                       N, ea, eb, js, C, ldc, pl.tb, jc)
        if (pl.tb.n <= 8) F4J_RECONSTRUCT(8);
        else if (pl.tb.n <= 12) F4J_RECONSTRUCT(12);
        else if (pl.tb.n <= 16) F4J_RECONSTRUCT(16);
        else F4J_RECONSTRUCT(CRT_MAX_MODULI);
#undef F4J_RECONSTRUCT
    }
    return hipGetLastError();
}

}  // namespace

size_t crt_job_stride(const CrtPlan &pl) { return (pl.bytes + 4095) & ~size_t(4095); }

hipError_t launch_crt_gemm(const CrtPlan &pl, const CrtOperand &oa, const CrtOperand &ob, int64_t M, int64_t N,
                           int64_t K, double *C, int64_t ldc, const CrtOutput &out, void *ws, hipStream_t stream,
                           hipEvent_t ev_begin, hipEvent_t ev_end, const CrtSide *side) {
    CrtCombos ccA{}, ccB{};
    ccA.ca[0] = oa.ca, ccA.cb[0] = oa.cb, ccB.ca[0] = ob.ca, ccB.cb[0] = ob.cb;
    return enqueue_crt<1>(pl, oa, ob, ccA, ccB, M, N, K, C, ldc, out, CrtJobCoef{}, ws, stream, ev_begin, ev_end,
                          side);
}

hipError_t launch_crt_gemm_jobs(const CrtPlan &pl, const CrtOperand &oa, const CrtOperand &ob, int64_t M, int64_t N,
                                int64_t K, double *C, int64_t ldc, const CrtJobs &jobs, void *ws, hipStream_t stream,
                                hipEvent_t ev_begin, hipEvent_t ev_end, const CrtSide *side) {
    CrtCombos ccA{}, ccB{};
    CrtJobCoef jc{};
    for (int j = 0; j < jobs.n && j < CRT_MAX_JOBS; ++j) {
        ccA.ca[j] = jobs.a_ca[j], ccA.cb[j] = jobs.a_cb[j];
        ccB.ca[j] = jobs.b_ca[j], ccB.cb[j] = jobs.b_cb[j];
        jc.cre[j] = jobs.cre[j], jc.cim[j] = jobs.cim[j];
    }
    CrtOutput out;
    out.cplx = true;
    out.first = true;
    switch (jobs.n) {
        case 3:
            return enqueue_crt<3>(pl, oa, ob, ccA, ccB, M, N, K, C, ldc, out, jc, ws, stream, ev_begin, ev_end, side);
        case 4:
            return enqueue_crt<4>(pl, oa, ob, ccA, ccB, M, N, K, C, ldc, out, jc, ws, stream, ev_begin, ev_end, side);
        default:
            return hipErrorInvalidValue;
    }
}

}  // namespace f4j
