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
