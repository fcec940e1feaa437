// Device building blocks of the split-precision (Ozaki-style) GEMM: the digit
// split of one FP64 value and the MFMA wrapper (shared by the single-product
// kernels in split_gemm.hip and the batched ones in split_batch.hip), and the k
// loop of one output tile over a list of digit pairs (segments) of the batched
// slice-product kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <type_traits>

namespace f4j {
namespace slice {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int BK = 64;                         // k slice (elements)
constexpr int LD = BK + 8;                     // LDS row stride in elements (16-byte pad)

// Workgroup tile of 2 x 2 waves, each wave BI x BJ MFMA blocks of 32 x 32.
template <int BI, int BJ>
struct Tile {
    static constexpr int BM = 64 * BI, BN = 64 * BJ;
    static constexpr int STAGE = (BM + BN) * LD;   // elements per LDS stage
};

// Signed radix-2^R digits of v 2^-e (|v| < 2^e): v 2^-e = 2^-(R-1) (d_0 + 2^-R (d_1 + ...)).
// Every step is exact in FP64; |d_s| <= 2^(R-1) fits the carrier T exactly.
template <int R, int S, class T>
__device__ inline void digits(double v, int e, T *d) {
    double y = (e == INT_MIN) ? 0.0 : ldexp(v, (R - 1) - e);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const double q = rint(y);
        d[s] = (T)(float)q;
        y = (y - q) * (double)(1 << R);
    }
}

// v_mfma_f32_32x32x16_{bf16,f16}: A/B lane l holds row l%32, k = 8(l/32)+v;
// result item r of lane l is row 8(r/4) + 4(l/32) + r%4, column l%32.
template <class T>
__device__ inline f32x16 mfma32(const u32x4 &a, const u32x4 &b, const f32x16 &c) {
    if constexpr (std::is_same<T, __bf16>::value)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                      0, 0);
}

// Row / column of accumulator item r of `lane` in wave block (i, j) of a tile
// whose wave owns rows wm.. and columns wn...
__device__ inline int acc_row(int wm, int i, int r, int lane) { return wm + 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
__device__ inline int acc_col(int wn, int j, int lane) { return wn + 32 * j + (lane & 31); }

// acc64 = sum over segments g of 2^shift(g) * sum_k A_{a(g)}[m0 + .][k] B_{b(g)}[n0 + .][k]
// for this wave's 32 BI x 32 BJ part of the tile (waves 2 x 2).  A and B are digit
// slices [S][Xp][Kp] (k contiguous, Kp a multiple of BK).  Digit products are
// at most 2^(2R-2), so the FP32 MFMA accumulator is exact over FLUSH
// consecutive k; it is added (scaled) to the FP64 accumulator every FLUSH k
// and at the end of each segment.  Seg provides a(g), b(g), shift(g).
// The caller's LDS buffer holds 2 * Tile<BI, BJ>::STAGE elements; the loop
// ends with a barrier, so the buffer may be reused right after.
template <class T, int FLUSH, int BI, int BJ, class Seg>
__device__ __forceinline__ void tile_loop(const void *A, const void *B, int64_t Mp, int64_t Np, int64_t Kp, int nseg,
                                          const Seg &seg, int64_t m0, int64_t n0, T *lds,
                                          double (&acc64)[BI][BJ][16]) {
    constexpr int TBM = Tile<BI, BJ>::BM, TSTAGE = Tile<BI, BJ>::STAGE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32 * BI, wn = (wave & 1) * 32 * BJ;
    f32x16 acc[BI][BJ];
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < BJ; ++j) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                acc[i][j][r] = 0.f;
                acc64[i][j][r] = 0.0;
            }
        }

    const int tps = (int)(Kp / BK);                  // k-tiles per segment
    const int ntiles = tps * nseg;
    constexpr int FLUSH_TILES = FLUSH / BK;
    // per-lane load geometry: passes of 32 rows x 8 16-byte columns (2 BI for A, 2 BJ for B)
    constexpr int PA = 2 * BI, PB = 2 * BJ, PM = PA > PB ? PA : PB;
    const int lrow = tid >> 3, lcol = (tid & 7) * 8;
    int lane_off[PM];
#pragma unroll
    for (int p = 0; p < PM; ++p) lane_off[p] = (int)(((int64_t)(lrow + 32 * p) * Kp + lcol) * (int64_t)sizeof(T));
    u32x4 ra[PA], rb[PB];

    // (segment, k offset) of the tile to load next, advanced without division
    int ld_seg = 0, ld_kt = 0;
    auto gload = [&]() {
        const int64_t kk = (int64_t)ld_kt * BK;
        const T *pa = (const T *)A + (int64_t)seg.a(ld_seg) * Mp * Kp + m0 * Kp + kk;
        const T *pb = (const T *)B + (int64_t)seg.b(ld_seg) * Np * Kp + n0 * Kp + kk;
        const __amdgpu_buffer_rsrc_t rsa = __builtin_amdgcn_make_buffer_rsrc((void *)pa, (short)0, 0x7FFFFFF0, 0x00020000);
        const __amdgpu_buffer_rsrc_t rsb = __builtin_amdgcn_make_buffer_rsrc((void *)pb, (short)0, 0x7FFFFFF0, 0x00020000);
#pragma unroll
        for (int p = 0; p < PA; ++p)
            ra[p] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rsa, lane_off[p], 0, 0));
#pragma unroll
        for (int p = 0; p < PB; ++p)
            rb[p] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rsb, lane_off[p], 0, 0));
        if (++ld_kt == tps) {
            ld_kt = 0;
            ++ld_seg;
        }
    };
    auto lstore = [&](T *st) {
#pragma unroll
        for (int p = 0; p < PA; ++p) *(u32x4 *)(st + (lrow + 32 * p) * LD + lcol) = ra[p];
#pragma unroll
        for (int p = 0; p < PB; ++p) *(u32x4 *)(st + TBM * LD + (lrow + 32 * p) * LD + lcol) = rb[p];
    };

    gload();
    lstore(lds);
    __syncthreads();
    const int fr = lane & 31, fk = (lane >> 5) * 8;
    int sg = 0, kt_in_seg = 0, since_flush = 0;
    for (int t = 0; t < ntiles; ++t) {
        const T *cur = lds + (t & 1) * TSTAGE;
        if (t + 1 < ntiles) gload();
#pragma unroll
        for (int ks = 0; ks < BK / 16; ++ks) {
            u32x4 fa[BI], fb[BJ];
#pragma unroll
            for (int i = 0; i < BI; ++i) fa[i] = *(const u32x4 *)(cur + (wm + 32 * i + fr) * LD + ks * 16 + fk);
#pragma unroll
            for (int j = 0; j < BJ; ++j) fb[j] = *(const u32x4 *)(cur + TBM * LD + (wn + 32 * j + fr) * LD + ks * 16 + fk);
#pragma unroll
            for (int i = 0; i < BI; ++i)
#pragma unroll
                for (int j = 0; j < BJ; ++j) acc[i][j] = mfma32<T>(fa[i], fb[j], acc[i][j]);
        }
        ++kt_in_seg;
        ++since_flush;
        if (since_flush == FLUSH_TILES || kt_in_seg == tps) {
            const double sc = __builtin_bit_cast(double, (uint64_t)(1023 + seg.shift(sg)) << 52);
#pragma unroll
            for (int i = 0; i < BI; ++i)
#pragma unroll
                for (int j = 0; j < BJ; ++j)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        acc64[i][j][r] = fma((double)acc[i][j][r], sc, acc64[i][j][r]);
                        acc[i][j][r] = 0.f;
                    }
            since_flush = 0;
            if (kt_in_seg == tps) {
                kt_in_seg = 0;
                ++sg;
            }
        }
        if (t + 1 < ntiles) lstore(lds + ((t + 1) & 1) * TSTAGE);
        __syncthreads();
    }
}

}  // namespace slice
}  // namespace f4j
