// Index math of the modular (CRT) GEMM, shared by its kernels and by the host
// emulator that checks it (tests/emu/crt_emu.cpp).
//
// Residue slices are stored fragment-major, in the order the INT8 MFMA reads
// them.  For a padded Xp x Kp slice (Xp, Kp multiples of 128), the 32 lines x
// 32 bytes block (x / 32, k / 32) is 1 KB, blocks ordered with k fastest
// (KC = Kp / 32 per line block); within a block, lane L = x % 32 + 32 ((k / 16) % 2)
// owns bytes 16 L .. 16 L + 15, byte k % 16.  That is the operand layout of
// v_mfma_i32_32x32x32_i8 (lane holds line lane % 32, k = 16 (lane / 32) + v),
// so one 16-byte load per lane reads a whole fragment from 1 KB of contiguous
// memory.
#pragma once
#include <cstdint>

#ifndef F4J_HD
#define F4J_HD __host__ __device__ inline
#endif

namespace f4j {
namespace crt {

constexpr int BM = 256, BN = 128, BK = 128;   // block tile (rows, columns); k slice in bytes
constexpr int FRAG = 1024;                    // bytes of one 32 x 32 fragment
constexpr int SLICE = (BK / 32) * FRAG;       // bytes of one line block in one k slice
constexpr int STAGES = 3;                     // B slices in flight through LDS
constexpr int STAGE = BN * BK;                // LDS bytes of one B stage (4 column blocks x 4 k steps)

F4J_HD int64_t frag_offset(int64_t x, int64_t k, int64_t KC) {
    return (((x >> 5) * KC + (k >> 5)) << 10) + (((x & 31) + ((k >> 4) & 1) * 32) << 4) + (k & 15);
}

// Split kernel (block = 32 lines x 128 k): thread t writes the 16 bytes of
// line x0 + split_line(t) starting at k0 + split_k(t).
F4J_HD int split_line(int t) { return t & 31; }
F4J_HD int split_k(int t) { return 16 * ((t >> 5) & 1) + 32 * (t >> 6); }

// Split kernel for strided lines (columns of a row-major operand; block = 128
// lines x 32 k, so each k row of the block is 1 KB of contiguous memory):
// thread t writes the 16 bytes of line x0 + split_cols_line(t) at k0 + split_cols_k(t).
F4J_HD int split_cols_line(int t) { return t & 127; }
F4J_HD int split_cols_k(int t) { return 16 * (t >> 7); }

// Product kernel.  Wave w of a block owns rows m0w = tile_m BM + 64 w .. + 63
// (register sets i = 0, 1 of 32 rows); the block's columns are n0 .. n0 + 127.
// A load for k slice t, set i, step ks: uniform base + soffset + per-lane voffset.
F4J_HD int64_t a_base(int64_t m0w, int t, int KC) { return (((m0w >> 5) * KC) << 10) + (int64_t)t * SLICE; }
F4J_HD int a_soff(int i, int KC) { return (i * KC) << 10; }
F4J_HD int a_voff(int ks, int lane) { return FRAG * ks + 16 * lane; }
// B: wave w loads column block w of slice t (4 fragments, p = k step) into
// LDS stage slot % STAGES; the hardware writes lane L's 16 bytes at +16 L.
F4J_HD int64_t b_base(int64_t n0, int t, int KC) { return (((n0 >> 5) * KC) << 10) + (int64_t)t * SLICE; }
F4J_HD int b_soff(int w, int p, int KC) { return ((w * KC) << 10) + FRAG * p; }
F4J_HD int b_voff(int lane) { return 16 * lane; }
F4J_HD int b_lds(int slot, int w, int p) { return (slot % STAGES) * STAGE + 4 * FRAG * w + FRAG * p; }
// Fragment read of column block j, k step ks, from stage slot.
F4J_HD int b_frag(int slot, int j, int ks, int lane) {
    return (slot % STAGES) * STAGE + 4 * FRAG * j + FRAG * ks + 16 * lane;
}
// MFMA result item r of `lane` in block (i, j): row and column within the wave tile.
F4J_HD int acc_row(int i, int r, int lane) { return 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
F4J_HD int acc_col(int j, int lane) { return 32 * j + (lane & 31); }

}  // namespace crt
}  // namespace f4j
