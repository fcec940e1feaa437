// FP64 / complex128 dense GEMM on CDNA4 FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// Replaces the Java kernels of RealDenseMatMult (M:linalg/ops/dense/real/
// RealDenseMatMult.java:50-415), RealDenseMatMultTranspose (.../real/
// RealDenseMatMultTranspose.java:53-233) and DenseSemiringMatMult{,Transpose}
// for Complex128 (M:linalg/ops/dense/semiring_ops/DenseSemiringMatMult.java:54-424,
// DenseSemiringMatMultTranspose.java:54-178), M: = src/main/java/org/flag4j/.
//
// One kernel template serves every operand kind.  A complex128 product is
// computed as the real product of its "real-expanded" operands (see OpKind in
// f4j_internal.h), so complex data runs on the same FP64 MFMA pipeline with
// the classical 8MNK flop count and no separate 4M code: the expansion happens
// in the per-lane LDS fragment addressing (a swap of re/im and a sign).
//
// Tiling: a workgroup owns a BM x BN tile of C~ and walks K~ in BK slices.
// Both operand slices are staged through LDS, double-buffered (one barrier per
// slice), with the next slice's global loads in flight while the current one
// feeds the MFMAs.  Each wave owns a (BM/WGM) x (BN/WGN) sub-tile held as
// 16x16 MFMA accumulators.  Loads are bounds-checked with zero fill, so any
// M, N, K (including ragged edges) is valid.
#include <type_traits>

#include "gemm_tile.h"

namespace f4j {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int WAVE = 64;

template <class O, int NT, int V>
struct Loader {
    using TC = TileCopy<O, NT, V>;
    typedef double dbl2 __attribute__((ext_vector_type(2)));
    using vec_t = typename std::conditional<V == 2, dbl2, double>::type;
    vec_t r[TC::PASSES];

    __device__ void load(const double *G, int64_t ld, int64_t x0, int64_t k0, int64_t X, int64_t K, int tid) {
        int64_t r0, c0, rows, cols;
        O::block(x0, k0, X, K, r0, c0, rows, cols);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(G + r0 * ld + c0), (short)0, (int)TC::num_records(r0, c0, rows, cols, ld), 0x00020000);
#pragma unroll
        for (int i = 0; i < TC::PASSES; ++i) {
            const uint32_t off = TC::offset(tid, i, c0, cols, ld);
            if constexpr (V == 2)
                r[i] = __builtin_bit_cast(vec_t, __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0));
            else
                r[i] = __builtin_bit_cast(vec_t, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off, 0, 0));
        }
    }
    __device__ void store(double *lds, int tid) const {
#pragma unroll
        for (int i = 0; i < TC::PASSES; ++i) {
            double *d = lds + TC::lds_index(tid, i);
            if constexpr (V == 2) {
                d[0] = r[i][0];
                d[1] = r[i][1];
            } else {
                d[0] = r[i];
            }
        }
    }
};

// One BM x BN output tile at (m0, n0) of problem p.  `lds` holds two stages.
// Every loop iteration ends in a barrier, so on return no wave still reads
// `lds` and the caller may start the next tile at once.
template <int BM, int BN, int BK, int WGM, int WGN, int V, int KA, int KB>
__device__ inline void gemm_tile(const GemmProblem &p, int64_t m0, int64_t n0, double *lds) {
    constexpr int NT = WGM * WGN * WAVE;
    constexpr int WTM = BM / WGM, WTN = BN / WGN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    using OA = Opnd<KA, true, BM, BK>;
    using OB = Opnd<KB, false, BN, BK>;
    constexpr int LDS_STAGE = OA::LDS_DOUBLES + OB::LDS_DOUBLES;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave / WGN) * WTM, wn = (wave % WGN) * WTN;
    const int baseA = OA::lane_base(lane, wm), baseB = OB::lane_base(lane, wn);
    const double sgnA = OA::lane_sign(lane), sgnB = OB::lane_sign(lane);

    dbl4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    Loader<OA, NT, V> la;
    Loader<OB, NT, V> lb;

    const int64_t nk = (p.K + BK - 1) / BK;
    la.load(p.A, p.lda, m0, 0, p.M, p.K, tid);
    lb.load(p.B, p.ldb, n0, 0, p.N, p.K, tid);
    la.store(lds, tid);
    lb.store(lds + OA::LDS_DOUBLES, tid);
    __syncthreads();

    for (int64_t kt = 0; kt < nk; ++kt) {
        const double *cur = lds + (kt & 1) * LDS_STAGE;
        if (kt + 1 < nk) {
            la.load(p.A, p.lda, m0, (kt + 1) * BK, p.M, p.K, tid);
            lb.load(p.B, p.ldb, n0, (kt + 1) * BK, p.N, p.K, tid);
        }
        const double *sa = cur, *sb = cur + OA::LDS_DOUBLES;
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            double a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                a[i] = sa[baseA + ks * OA::KS + i * OA::FS];
                if constexpr (OA::cplx) a[i] *= sgnA;
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                b[j] = sb[baseB + ks * OB::KS + j * OB::FS];
                if constexpr (OB::cplx) b[j] *= sgnB;
            }
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < nk) {
            double *nxt = lds + ((kt + 1) & 1) * LDS_STAGE;
            la.store(nxt, tid);
            lb.store(nxt + OA::LDS_DOUBLES, tid);
        }
        __syncthreads();
    }

    const int64_t col_l = n0 + wn + mfma_f64_acc_col(lane);
    const int64_t row_l = m0 + wm;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int64_t col = col_l + j * 16;
            if (col >= p.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = row_l + i * 16 + mfma_f64_acc_row(lane, r);
                if (row < p.M) p.C[row * p.ldc + col] = acc[i][j][r];
            }
        }
}

template <int BM, int BN, int BK, int WGM, int WGN, int MINW, int V, int KA, int KB>
__global__ __launch_bounds__(WGM *WGN *WAVE, MINW) void gemm_f64_kernel(GemmProblem p) {
    constexpr int LDS_STAGE = Opnd<KA, true, BM, BK>::LDS_DOUBLES + Opnd<KB, false, BN, BK>::LDS_DOUBLES;
    __shared__ double lds[2 * LDS_STAGE];
    const int64_t tiles_m = (p.M + BM - 1) / BM, tiles_n = (p.N + BN - 1) / BN;
    int64_t tm, tn;
    tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    gemm_tile<BM, BN, BK, WGM, WGN, V, KA, KB>(p, tm * BM, tn * BN, lds);
}

// Persistent grouped kernel: workgroup g computes the work items
// items[begin[g] .. begin[g+1]), each one tile of one problem.  Problems whose
// rows are 16-byte aligned (flags bit 0) use 16-byte loads.  With guard.err
// set, only the items whose problem the guard selects run (uniform per item).
template <int BM, int BN, int BK, int WGM, int WGN, int MINW, int KA, int KB>
__global__ __launch_bounds__(WGM *WGN *WAVE, MINW) void gemm_f64_grouped_kernel(const GemmProblem *probs,
                                                                             const GroupedItem *items,
                                                                             const int *begin, GroupedGuard guard) {
    constexpr int LDS_STAGE = Opnd<KA, true, BM, BK>::LDS_DOUBLES + Opnd<KB, false, BN, BK>::LDS_DOUBLES;
    __shared__ double lds[2 * LDS_STAGE];
    const int b0 = begin[blockIdx.x], b1 = begin[blockIdx.x + 1];
    for (int it = b0; it < b1; ++it) {
        const GroupedItem w = items[it];
        if (guard.err) {
            const int s = guard.slot[w.prob];
            if (s < 0 || !(guard.err[s] > guard.above)) continue;
        }
        const GemmProblem p = probs[w.prob];
        if (p.flags & 1)
            gemm_tile<BM, BN, BK, WGM, WGN, 2, KA, KB>(p, (int64_t)w.tm * BM, (int64_t)w.tn * BN, lds);
        else
            gemm_tile<BM, BN, BK, WGM, WGN, 1, KA, KB>(p, (int64_t)w.tm * BM, (int64_t)w.tn * BN, lds);
    }
}

// Direct-load grouped kernel for real operands (tile configurations
// F4J_DIRECT_BASE + i): one wave per (16 FM) x (16 FN) tile, fragments read
// straight from global memory into registers -- no LDS, no barrier, so waves
// of one workgroup never wait for each other and short k-loops cost no
// per-slice synchronisation.  For k step s (k = 4 s .. 4 s + 3) lane l holds
// A(m0 + 16 i + l % 16, 4 s + l / 16) and B(4 s + l / 16, n0 + 16 j + l % 16),
// the operand layout of v_mfma_f64_16x16x4f64, so every element of C
// accumulates the same MFMA k steps in the same order, from zero, as in the
// LDS kernels: the results are the same bits.  Elements past M, N or K read a
// clamped in-range address and are replaced by zero (no branch around a load,
// so the D k steps of loads in flight are counted exactly).
template <int FM, int FN, int KA, int KB>
__global__ __launch_bounds__(64) void gemm_f64_direct_kernel(const GemmProblem *probs, const GroupedItem *items,
                                                             const int *begin, GroupedGuard guard) {
    static_assert(KA == R_N || KA == R_T, "real A");
    static_assert(KB == R_N || KB == R_T, "real B");
    constexpr int D = 4;                           // k steps of loads in flight
    const int lane = threadIdx.x, r = lane & 15, q = lane >> 4;
    const int b0 = begin[blockIdx.x], b1 = begin[blockIdx.x + 1];
    for (int it = b0; it < b1; ++it) {
        // the item is the wave's: made uniform explicitly, so the buffer
        // resources built from its problem stay in scalar registers
        const int prob = __builtin_amdgcn_readfirstlane(items[it].prob);
        const int tm = __builtin_amdgcn_readfirstlane(items[it].tm), tn = __builtin_amdgcn_readfirstlane(items[it].tn);
        if (guard.err) {
            const int sl = guard.slot[prob];
            if (sl < 0 || !(guard.err[sl] > guard.above)) continue;
        }
        GemmProblem p = probs[prob];
        auto uni = [](int64_t v) {
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
            return (int64_t)(((uint64_t)hi << 32) | lo);
        };
        p.M = uni(p.M), p.N = uni(p.N), p.K = uni(p.K), p.lda = uni(p.lda), p.ldb = uni(p.ldb), p.ldc = uni(p.ldc);
        p.A = (const double *)uni((int64_t)p.A), p.B = (const double *)uni((int64_t)p.B), p.C = (double *)uni((int64_t)p.C);
        const int64_t m0 = (int64_t)tm * 16 * FM, n0 = (int64_t)tn * 16 * FN;
        constexpr bool a_rows = KA == R_N, b_rows = KB == R_N;   // k contiguous in A / k the row index of B
        // one buffer resource per operand (extents below 2^31 bytes, checked when the plan is made);
        // per fragment a 32-bit byte offset of this lane's line, plus k times the k stride
        const int64_t ea = a_rows ? (p.M - 1) * p.lda + p.K : (p.K - 1) * p.lda + p.M;
        const int64_t eb = b_rows ? (p.K - 1) * p.ldb + p.N : (p.N - 1) * p.ldb + p.K;
        const __amdgpu_buffer_rsrc_t rsa = __builtin_amdgcn_make_buffer_rsrc((void *)p.A, (short)0, (int)(ea * 8),
                                                                             0x00020000);
        const __amdgpu_buffer_rsrc_t rsb = __builtin_amdgcn_make_buffer_rsrc((void *)p.B, (short)0, (int)(eb * 8),
                                                                             0x00020000);
        const uint32_t ka = a_rows ? 8u : (uint32_t)p.lda * 8u, kb = b_rows ? (uint32_t)p.ldb * 8u : 8u;
        uint32_t oa[FM], ob[FN];
        bool oka[FM], okb[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int64_t row = m0 + 16 * i + r;
            oka[i] = row < p.M;
            const uint32_t rc = (uint32_t)(oka[i] ? row : p.M - 1);
            oa[i] = a_rows ? rc * (uint32_t)p.lda * 8u : rc * 8u;
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int64_t col = n0 + 16 * j + r;
            okb[j] = col < p.N;
            const uint32_t cc = (uint32_t)(okb[j] ? col : p.N - 1);
            ob[j] = b_rows ? cc * 8u : cc * (uint32_t)p.ldb * 8u;
        }
        const int64_t nks = (p.K + 3) / 4;
        const uint32_t klast = (uint32_t)(p.K - 1);
        double ra[D][FM], rb[D][FN];
        // raw loads into the ring; the zeros for elements past M, N or K are
        // applied when the MFMAs consume them, so no load waits where it is issued
        auto fetch = [&](int64_t s, double (&a)[FM], double (&b)[FN]) {
            const int64_t k = 4 * s + q;
            const uint32_t kc = k < p.K ? (uint32_t)k : klast;
#pragma unroll
            for (int i = 0; i < FM; ++i)
                a[i] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsa, oa[i] + kc * ka, 0, 0));
#pragma unroll
            for (int j = 0; j < FN; ++j)
                b[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsb, ob[j] + kc * kb, 0, 0));
        };
        dbl4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
        auto mfmas = [&](int64_t s, const double (&a)[FM], const double (&b)[FN]) {
            const bool kin = 4 * s + q < p.K;
            double x[FM], y[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) x[i] = (kin && oka[i]) ? a[i] : 0.0;
#pragma unroll
            for (int j = 0; j < FN; ++j) y[j] = (kin && okb[j]) ? b[j] : 0.0;
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[i], y[j], acc[i][j], 0, 0, 0);
        };
        // K = 0: no loads at all (the operands may hold no element); C is zero
        if (nks > 0) {
#pragma unroll
            for (int d = 0; d < D; ++d) fetch(d, ra[d], rb[d]);
        }
        int64_t s = 0;
        for (; s + D <= nks; s += D) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                mfmas(s + d, ra[d], rb[d]);
                fetch(s + d + D, ra[d], rb[d]);       // past the last k step: clamped reloads, never used
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (s + d < nks) mfmas(s + d, ra[d], rb[d]);
        const int64_t col_l = n0 + mfma_f64_acc_col(lane);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int64_t col = col_l + 16 * j;
                if (col >= p.N) continue;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t row = m0 + 16 * i + mfma_f64_acc_row(lane, e);
                    if (row < p.M) p.C[row * p.ldc + col] = acc[i][j][e];
                }
            }
    }
}

// (FM, FN) of the direct configurations F4J_DIRECT_BASE + i.
#define F4J_DIRECT_TILES(X) X(0, 2, 2) X(1, 4, 2) X(2, 2, 4) X(3, 4, 4)

template <int KA, int KB>
hipError_t launch_direct_kinds(const GemmProblem *probs, const GroupedItem *items, const int *begin, int wgs, int cfg,
                               hipStream_t stream, GroupedGuard guard) {
    if constexpr ((KA == R_N || KA == R_T) && (KB == R_N || KB == R_T)) {
        switch (cfg - F4J_DIRECT_BASE) {
#define F4J_DIRECT_CASE(i, FM, FN)                                                                                  case i:                                                                                                            hipLaunchKernelGGL((gemm_f64_direct_kernel<FM, FN, KA, KB>), dim3((unsigned)wgs), dim3(64), 0, stream,                            probs, items, begin, guard);                                                                return hipGetLastError();
            F4J_DIRECT_TILES(F4J_DIRECT_CASE)
#undef F4J_DIRECT_CASE
            default: return hipErrorInvalidValue;
        }
    } else {
        (void)probs, (void)items, (void)begin, (void)wgs, (void)cfg, (void)stream, (void)guard;
        return hipErrorInvalidValue;          // complex operands: the LDS kernels only
    }
}

template <int KA, int KB>
int direct_occ_kinds(int cfg) {
    int n = 0;
    if constexpr ((KA == R_N || KA == R_T) && (KB == R_N || KB == R_T)) {
        switch (cfg - F4J_DIRECT_BASE) {
#define F4J_DIRECT_CASE(i, FM, FN)                                                                                  case i:                                                                                                            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_f64_direct_kernel<FM, FN, KA, KB>, 64, 0) !=             hipSuccess)                                                                                                    n = 1;                                                                                                     return n;
            F4J_DIRECT_TILES(F4J_DIRECT_CASE)
#undef F4J_DIRECT_CASE
            default: return 1;
        }
    }
    (void)cfg;
    return 1;
}

template <int BM, int BN, int BK, int WGM, int WGN, int MINW, int KA, int KB>
hipError_t launch_cfg(const GemmProblem &p, bool vec2, hipStream_t stream) {
    const int64_t tiles = ((p.M + BM - 1) / BM) * ((p.N + BN - 1) / BN);
    if (tiles == 0) return hipSuccess;
    if (vec2)
        hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, BK, WGM, WGN, MINW, 2, KA, KB>), dim3((unsigned)tiles),
                           dim3(WGM * WGN * WAVE), 0, stream, p);
    else
        hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, BK, WGM, WGN, MINW, 1, KA, KB>), dim3((unsigned)tiles),
                           dim3(WGM * WGN * WAVE), 0, stream, p);
    return hipGetLastError();
}

template <int KA, int KB>
hipError_t launch_kinds(const GemmProblem &p, bool vec2, hipStream_t stream, int cfg) {
    switch (cfg) {
#define F4J_TILE_CASE(id, BM, BN, BK, WGM, WGN, MINW) \
    case id: return launch_cfg<BM, BN, BK, WGM, WGN, MINW, KA, KB>(p, vec2, stream);
        F4J_FP64_TILES(F4J_TILE_CASE)
#undef F4J_TILE_CASE
        default: return hipErrorInvalidValue;
    }
}

template <int KA, int KB>
hipError_t launch_grouped_kinds(const GemmProblem *probs, const GroupedItem *items, const int *begin, int wgs, int cfg,
                                hipStream_t stream, GroupedGuard guard) {
    switch (cfg) {
#define F4J_TILE_CASE(id, BM, BN, BK, WGM, WGN, MINW)                                                        \
    case id:                                                                                                \
        hipLaunchKernelGGL((gemm_f64_grouped_kernel<BM, BN, BK, WGM, WGN, MINW, KA, KB>), dim3((unsigned)wgs), \
                           dim3(WGM * WGN * WAVE), 0, stream, probs, items, begin, guard);                   \
        return hipGetLastError();
        F4J_FP64_TILES(F4J_TILE_CASE)
#undef F4J_TILE_CASE
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

namespace {
template <int KA, int KB>
int grouped_occ_kinds(int cfg) {
    int n = 0;
    switch (cfg) {
#define F4J_TILE_CASE(id, BM, BN, BK, WGM, WGN, MINW)                                                           \
    case id:                                                                                                   \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm_f64_grouped_kernel<BM, BN, BK, WGM, WGN, MINW, KA, KB>, \
                                                         WGM * WGN * WAVE, 0) != hipSuccess)                   \
            n = 1;                                                                                             \
        return n;
        F4J_FP64_TILES(F4J_TILE_CASE)
#undef F4J_TILE_CASE
        default: return 1;
    }
}
}  // namespace

int grouped_wgs_per_cu(OpKind ka, OpKind kb, int cfg) {
    if (cfg >= F4J_DIRECT_BASE) {
        if (ka == R_N && kb == R_N) return direct_occ_kinds<R_N, R_N>(cfg);
        if (ka == R_N && kb == R_T) return direct_occ_kinds<R_N, R_T>(cfg);
        if (ka == R_T && kb == R_N) return direct_occ_kinds<R_T, R_N>(cfg);
        if (ka == R_T && kb == R_T) return direct_occ_kinds<R_T, R_T>(cfg);
        return 1;
    }
#define F4J_CASE(A_, B_) \
    if (ka == A_ && kb == B_) return grouped_occ_kinds<A_, B_>(cfg);
    F4J_CASE(R_N, R_N) F4J_CASE(R_N, R_T) F4J_CASE(R_T, R_N) F4J_CASE(R_T, R_T)
    F4J_CASE(Z_N, Z_N) F4J_CASE(Z_N, Z_T) F4J_CASE(Z_N, Z_C)
    F4J_CASE(Z_T, Z_N) F4J_CASE(Z_T, Z_T) F4J_CASE(Z_T, Z_C)
    F4J_CASE(Z_C, Z_N) F4J_CASE(Z_C, Z_T) F4J_CASE(Z_C, Z_C)
#undef F4J_CASE
    return 1;
}

hipError_t launch_gemm_f64_grouped(OpKind ka, OpKind kb, const GemmProblem *probs, const GroupedItem *items,
                                   const int *begin, int wgs, int cfg, hipStream_t stream, GroupedGuard guard) {
    if (wgs <= 0) return hipSuccess;
    if (cfg >= F4J_DIRECT_BASE) {
        if (ka == R_N && kb == R_N) return launch_direct_kinds<R_N, R_N>(probs, items, begin, wgs, cfg, stream, guard);
        if (ka == R_N && kb == R_T) return launch_direct_kinds<R_N, R_T>(probs, items, begin, wgs, cfg, stream, guard);
        if (ka == R_T && kb == R_N) return launch_direct_kinds<R_T, R_N>(probs, items, begin, wgs, cfg, stream, guard);
        if (ka == R_T && kb == R_T) return launch_direct_kinds<R_T, R_T>(probs, items, begin, wgs, cfg, stream, guard);
        return hipErrorInvalidValue;
    }
#define F4J_CASE(A_, B_) \
    if (ka == A_ && kb == B_) return launch_grouped_kinds<A_, B_>(probs, items, begin, wgs, cfg, stream, guard);
    F4J_CASE(R_N, R_N) F4J_CASE(R_N, R_T) F4J_CASE(R_T, R_N) F4J_CASE(R_T, R_T)
    F4J_CASE(Z_N, Z_N) F4J_CASE(Z_N, Z_T) F4J_CASE(Z_N, Z_C)
    F4J_CASE(Z_T, Z_N) F4J_CASE(Z_T, Z_T) F4J_CASE(Z_T, Z_C)
    F4J_CASE(Z_C, Z_N) F4J_CASE(Z_C, Z_T) F4J_CASE(Z_C, Z_C)
#undef F4J_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_gemm_f64(OpKind ka, OpKind kb, const GemmProblem &p, hipStream_t stream, int cfg) {
    const bool vec2 = vec2_ok(p.A, p.lda, stored_cols(ka, true, p.M, p.N, p.K)) &&
                      vec2_ok(p.B, p.ldb, stored_cols(kb, false, p.M, p.N, p.K));
#define F4J_CASE(A_, B_) \
    if (ka == A_ && kb == B_) return launch_kinds<A_, B_>(p, vec2, stream, cfg);
    F4J_CASE(R_N, R_N) F4J_CASE(R_N, R_T) F4J_CASE(R_T, R_N) F4J_CASE(R_T, R_T)
    F4J_CASE(Z_N, Z_N) F4J_CASE(Z_N, Z_T) F4J_CASE(Z_N, Z_C)
    F4J_CASE(Z_T, Z_N) F4J_CASE(Z_T, Z_T) F4J_CASE(Z_T, Z_C)
    F4J_CASE(Z_C, Z_N) F4J_CASE(Z_C, Z_T) F4J_CASE(Z_C, Z_C)
#undef F4J_CASE
    return hipErrorInvalidValue;
}

}  // namespace f4j
