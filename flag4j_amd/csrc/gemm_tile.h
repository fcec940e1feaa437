// Tile geometry and LDS fragment addressing of the FP64 MFMA GEMM, shared by
// the device kernel and the host-side emulator that checks the index math.
#pragma once
#include <cstdint>

#include "f4j_internal.h"

namespace f4j {

#define F4J_HD __host__ __device__ inline

// Geometry of one operand's stored tile and its LDS fragment addressing.
// X = m for A, n~ for B.  BX = tile extent along X, BK = along k~.
template <int KIND, bool IS_A, int BX, int BK>
struct Opnd {
    static constexpr bool cplx = KIND >= Z_N;
    // Stored block is laid out with rows over X ("x-major") or rows over k~.
    static constexpr bool x_major = IS_A ? (KIND == R_N || KIND == Z_N) : (KIND == R_T || KIND == Z_T || KIND == Z_C);
    // In the real view of complex storage one stored row covers two X values
    // (B transposed: a row is one n holding (re,im) pairs along k), two k~
    // values (A transposed, B not transposed), or its columns hold 2 values per
    // X (A transposed: (re,im) pairs along m).
    static constexpr bool xrows_half = cplx && !IS_A && x_major;
    static constexpr bool krows_half = cplx && !x_major;
    static constexpr bool xcols_double = cplx && IS_A && !x_major;
    // stored block shape (rows SR x cols SC) of the real view of the storage
    static constexpr int SR = x_major ? (xrows_half ? BX / 2 : BX) : (krows_half ? BK / 2 : BK);
    static constexpr int SC = x_major ? BK : (xcols_double ? 2 * BX : BX);
    // x-major tiles are read down a column by 16 lanes: an odd row stride
    // keeps those reads on distinct banks.
    static constexpr int PAD = x_major ? 1 : 0;
    static constexpr int LD = SC + PAD;
    static constexpr int LDS_DOUBLES = SR * LD;

    // Strides of the fragment index per k-step (4 k~) and per 16-wide fragment.
    static constexpr int KS = x_major ? 4 : (krows_half ? 2 * LD : 4 * LD);
    static constexpr int FS = x_major ? (xrows_half ? 8 * LD : 16 * LD) : (xcols_double ? 32 : 16);

    // LDS index of fragment element for lane `lane` at k-step 0, fragment 0,
    // wave offset `woff` along X.  MFMA 16x16x4 operand layout: lane l holds
    // X = l % 16 and k~ = l / 16.
    F4J_HD static int lane_base(int lane, int woff) {
        const int r = lane & 15, q = lane >> 4;
        const int s = q & 1;
        if constexpr (!cplx) {
            return x_major ? (woff + r) * LD + q : q * LD + woff + r;
        } else if constexpr (IS_A) {
            if constexpr (x_major)                 // Z_N: storage row m holds (re,im) pairs along k
                return (woff + r) * LD + q;
            else                                   // Z_T / Z_C: storage row k holds (re,im) pairs along m
                return (q >> 1) * LD + 2 * (woff + r) + s;
        } else {
            if constexpr (!x_major)                // Z_N: storage row k holds (re,im) pairs along n
                return (q >> 1) * LD + ((woff + r) ^ s);
            else                                   // Z_T / Z_C: storage row n holds (re,im) pairs along k
                return ((woff + r) >> 1) * LD + ((q & ~1) | ((r & 1) ^ s));
        }
    }

    // Sign applied to the fragment element (real expansion and conjugation).
    F4J_HD static double lane_sign(int lane) {
        const int s = (lane >> 4) & 1, c = lane & 1;
        if constexpr (IS_A) return (KIND == Z_C && s) ? -1.0 : 1.0;
        if constexpr (KIND == Z_N || KIND == Z_T) return (s && !c) ? -1.0 : 1.0;
        if constexpr (KIND == Z_C) return (!s && c) ? -1.0 : 1.0;
        return 1.0;
    }

    // Origin of the stored block for tile (x0, k0) in real-expanded units,
    // and the extents of the stored real matrix.
    F4J_HD static void block(int64_t x0, int64_t k0, int64_t X, int64_t K, int64_t &r0, int64_t &c0, int64_t &rows,
                             int64_t &cols) {
        if constexpr (x_major) {
            r0 = xrows_half ? x0 / 2 : x0;    rows = xrows_half ? X / 2 : X;
            c0 = k0;                          cols = K;
        } else {
            r0 = krows_half ? k0 / 2 : k0;    rows = krows_half ? K / 2 : K;
            c0 = xcols_double ? 2 * x0 : x0;  cols = xcols_double ? 2 * X : X;
        }
    }

    // Element e of the stored block: row/col within the block and LDS index.
    F4J_HD static int e_row(int e) { return e / SC; }
    F4J_HD static int e_col(int e) { return e % SC; }
    F4J_HD static int e_lds(int e) { return (e / SC) * LD + e % SC; }
};

// Global -> LDS copy of one operand's stored block by NT threads with V-double
// vectors (V = 1 or 2), through a buffer resource whose base is the block's
// origin and whose num_records ends at the stored matrix's last element.
// Rows past the matrix end therefore fall outside num_records; columns past
// the matrix edge (which alias the next row in memory) get an offset beyond
// any num_records.  The hardware returns 0 for both, so edge tiles need no
// branches.  Every lane keeps the same column for all its passes.
template <class O, int NT, int V>
struct TileCopy {
    static constexpr int SCV = O::SC / V;             // vectors per stored row
    static constexpr int ROWS_PER_PASS = NT / SCV;
    static constexpr int PASSES = O::SR * SCV / NT;
    static_assert(NT % SCV == 0 && (O::SR * SCV) % NT == 0, "tile copy does not divide evenly");
    static constexpr uint32_t OOB = 0x80000000u;      // beyond every num_records used

    F4J_HD static int lane_col(int tid) { return V * (tid % SCV); }
    F4J_HD static int lane_row(int tid) { return tid / SCV; }
    F4J_HD static int lds_index(int tid, int pass) { return (lane_row(tid) + pass * ROWS_PER_PASS) * O::LD + lane_col(tid); }
    // num_records (bytes) for a block at (r0, c0) of a rows x cols matrix; 0 for a
    // block that starts past the matrix's last element (e.g. K = 0: no rows at all),
    // so that every load of it returns zero.
    F4J_HD static uint32_t num_records(int64_t r0, int64_t c0, int64_t rows, int64_t cols, int64_t ld) {
        const int64_t n = ((rows - 1 - r0) * ld + (cols - c0)) * 8;
        return n <= 0 ? 0u : n > (int64_t)0x7FFFFFF0 ? 0x7FFFFFF0u : (uint32_t)n;
    }
    // Byte offset of pass `pass` for lane `tid`, or OOB if its columns are past the edge.
    F4J_HD static uint32_t offset(int tid, int pass, int64_t c0, int64_t cols, int64_t ld) {
        const int c = lane_col(tid);
        if (c0 + c + V > cols) return OOB;
        return (uint32_t)(((int64_t)(lane_row(tid) + pass * ROWS_PER_PASS) * ld + c) * 8);
    }
};

// Row length of an operand's stored real view (real-expanded sizes M, N, K).
F4J_HD int64_t stored_cols(int kind, bool is_a, int64_t M, int64_t N, int64_t K) {
    const bool trans = (kind == R_T || kind == Z_T || kind == Z_C);
    if (is_a) return trans ? (kind >= Z_N ? 2 * M : M) : K;
    return trans ? K : N;
}

// 16-byte vector loads need 16-byte aligned rows and an even stored row
// length, so that a vector never straddles the matrix edge.
F4J_HD bool vec2_ok(const double *ptr, int64_t ld, int64_t cols) {
    return ((uintptr_t)ptr % 16 == 0) && ld % 2 == 0 && cols % 2 == 0;
}

// Result layout of v_mfma_f64_16x16x4f64: lane l, item r of the 4-double
// accumulator holds D[l/16 + 4r][l%16] (rows interleave across the 4 items;
// the FP32/BF16 16x16 forms instead hold 4 consecutive rows per lane).
F4J_HD int mfma_f64_acc_row(int lane, int item) { return (lane >> 4) + 4 * item; }
F4J_HD int mfma_f64_acc_col(int lane) { return lane & 15; }

// FP64 MFMA tile configurations: X(id, BM, BN, BK, WGM, WGN, MINW) with a
// BM x BN workgroup tile walked in BK slices by WGM x WGN waves and at least
// MINW waves resident per SIMD.  Shared by the launcher, the dispatcher's
// names and the host emulator.
#define F4J_FP64_TILES(X)            \
    X(0, 128, 128, 16, 2, 2, 2)      \
    X(1, 64, 64, 16, 2, 2, 2)        \
    X(2, 128, 128, 32, 2, 2, 1)      \
    X(3, 128, 128, 16, 2, 2, 1)      \
    X(4, 128, 128, 16, 4, 2, 4)      \
    X(5, 128, 128, 32, 4, 2, 2)      \
    X(6, 256, 128, 16, 4, 2, 2)      \
    X(7, 64, 128, 16, 2, 2, 4)       \
    X(8, 128, 64, 16, 2, 2, 4)

#define F4J_COUNT_TILE(id, BM, BN, BK, WGM, WGN, MINW) +1
constexpr int NUM_FP64_TILES = 0 F4J_FP64_TILES(F4J_COUNT_TILE);
#undef F4J_COUNT_TILE

// Workgroup id -> (tile row, tile col).  The hardware deals consecutive
// workgroup ids round-robin over the 8 XCDs; remap so each XCD walks a
// contiguous band of tiles, then order the band in groups of GROUP tile-rows
// so concurrently running tiles share B panels in the XCD's L2.
F4J_HD void tile_of(int64_t bid, int64_t tiles_m, int64_t tiles_n, int64_t &tm, int64_t &tn) {
    constexpr int GROUP = 8;
    const int64_t ntiles = tiles_m * tiles_n;
    if (ntiles % 8 == 0) bid = (bid % 8) * (ntiles / 8) + bid / 8;
    const int64_t per_group = GROUP * tiles_n;
    const int64_t g = bid / per_group;
    const int64_t first_m = g * GROUP;
    const int64_t gsize = (tiles_m - first_m) < GROUP ? (tiles_m - first_m) : GROUP;
    tm = first_m + (bid % per_group) % gsize;
    tn = (bid % per_group) / gsize;
}

}  // namespace f4j
