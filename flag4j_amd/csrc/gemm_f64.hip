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
#include "gemm_tile.h"

namespace f4j {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int WAVE = 64;

template <int BM, int BN, int BK, int WGM, int WGN, int KA, int KB>
__global__ __launch_bounds__(WGM *WGN *WAVE) void gemm_f64_kernel(GemmProblem p) {
    constexpr int NT = WGM * WGN * WAVE;
    constexpr int WTM = BM / WGM, WTN = BN / WGN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    using OA = Opnd<KA, true, BM, BK>;
    using OB = Opnd<KB, false, BN, BK>;
    static_assert((OA::SR * OA::SC) % NT == 0 && (OB::SR * OB::SC) % NT == 0, "tile/threads mismatch");
    constexpr int EA = OA::SR * OA::SC / NT, EB = OB::SR * OB::SC / NT;
    constexpr int LDS_STAGE = OA::LDS_DOUBLES + OB::LDS_DOUBLES;

    __shared__ double lds[2 * LDS_STAGE];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    const int64_t tiles_m = (p.M + BM - 1) / BM, tiles_n = (p.N + BN - 1) / BN;
    int64_t tm, tn;
    tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
    const int64_t m0 = tm * BM, n0 = tn * BN;

    const int wm = (wave / WGN) * WTM, wn = (wave % WGN) * WTN;
    const int baseA = OA::lane_base(lane, wm), baseB = OB::lane_base(lane, wn);
    const double sgnA = OA::lane_sign(lane), sgnB = OB::lane_sign(lane);

    dbl4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    double ra[EA], rb[EB];

    auto gload = [&](int64_t k0) {
        int64_t r0, c0, rows, cols;
        OA::block(m0, k0, p.M, p.K, r0, c0, rows, cols);
#pragma unroll
        for (int i = 0; i < EA; ++i) {
            const int e = tid + i * NT;
            const int64_t r = r0 + OA::e_row(e), c = c0 + OA::e_col(e);
            ra[i] = (r < rows && c < cols) ? p.A[r * p.lda + c] : 0.0;
        }
        OB::block(n0, k0, p.N, p.K, r0, c0, rows, cols);
#pragma unroll
        for (int i = 0; i < EB; ++i) {
            const int e = tid + i * NT;
            const int64_t r = r0 + OB::e_row(e), c = c0 + OB::e_col(e);
            rb[i] = (r < rows && c < cols) ? p.B[r * p.ldb + c] : 0.0;
        }
    };
    auto lstore = [&](double *stage) {
#pragma unroll
        for (int i = 0; i < EA; ++i) stage[OA::e_lds(tid + i * NT)] = ra[i];
        double *sb = stage + OA::LDS_DOUBLES;
#pragma unroll
        for (int i = 0; i < EB; ++i) sb[OB::e_lds(tid + i * NT)] = rb[i];
    };

    const int64_t nk = (p.K + BK - 1) / BK;
    gload(0);
    lstore(lds);
    __syncthreads();

    for (int64_t kt = 0; kt < nk; ++kt) {
        const double *cur = lds + (kt & 1) * LDS_STAGE;
        if (kt + 1 < nk) gload((kt + 1) * BK);
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
        if (kt + 1 < nk) lstore(lds + ((kt + 1) & 1) * LDS_STAGE);
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

template <int BM, int BN, int BK, int WGM, int WGN, int KA, int KB>
hipError_t launch_cfg(const GemmProblem &p, hipStream_t stream) {
    const int64_t tiles = ((p.M + BM - 1) / BM) * ((p.N + BN - 1) / BN);
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, BK, WGM, WGN, KA, KB>), dim3((unsigned)tiles),
                       dim3(WGM * WGN * WAVE), 0, stream, p);
    return hipGetLastError();
}

template <int KA, int KB>
hipError_t launch_kinds(const GemmProblem &p, hipStream_t stream, int cfg) {
    switch (cfg) {
        case 1: return launch_cfg<64, 64, 16, 2, 2, KA, KB>(p, stream);
        case 2: return launch_cfg<128, 128, 32, 2, 2, KA, KB>(p, stream);
        default: return launch_cfg<128, 128, 16, 2, 2, KA, KB>(p, stream);
    }
}

}  // namespace

hipError_t launch_gemm_f64(OpKind ka, OpKind kb, const GemmProblem &p, hipStream_t stream, int cfg) {
#define F4J_CASE(A_, B_) \
    if (ka == A_ && kb == B_) return launch_kinds<A_, B_>(p, stream, cfg);
    F4J_CASE(R_N, R_N) F4J_CASE(R_N, R_T) F4J_CASE(R_T, R_N) F4J_CASE(R_T, R_T)
    F4J_CASE(Z_N, Z_N) F4J_CASE(Z_N, Z_T) F4J_CASE(Z_N, Z_C)
    F4J_CASE(Z_T, Z_N) F4J_CASE(Z_T, Z_T) F4J_CASE(Z_T, Z_C)
    F4J_CASE(Z_C, Z_N) F4J_CASE(Z_C, Z_T) F4J_CASE(Z_C, Z_C)
#undef F4J_CASE
    return hipErrorInvalidValue;
}

}  // namespace f4j
