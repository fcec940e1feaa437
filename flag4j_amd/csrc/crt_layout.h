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

constexpr int RB = 2;                         // 32-row line blocks per wave (wave tile 32 RB x 128)
constexpr int BM = 4 * 32 * RB, BN = 256;     // product block tile (rows, columns)
constexpr int BK = 128;                       // k padding unit of the residue slices, bytes
constexpr int FRAG = 1024;                    // bytes of one 32 x 32 fragment
constexpr int WAVES = 8;                      // product block: 4 row groups of 32 RB x 2 column halves of 128
constexpr int SK = 128;                       // k bytes per product step
constexpr int KB = SK / 32;                   // 32-byte MFMA k steps (k blocks) per product step
constexpr int STAGES = 2;                     // steps in the LDS ring
constexpr int AHEAD = STAGES - 1;             // steps in flight ahead of the MFMAs
constexpr int LOADS_A = RB * KB;              // fragment loads per step of an A wave (waves 0-3)
constexpr int LOADS_B = 2 * KB;               // and of a B wave (waves 4-7)
constexpr int STAGE = (BM + BN) * SK;         // LDS bytes of one step: its A fragments, then its B fragments

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

// Product kernel.  A block computes a BM x 256 tile.  Every step t (k bytes
// SK t .. SK t + SK - 1) both operands' fragments go straight from the residue
// slices into LDS stage t % STAGES, each fragment loaded once per block:
// stage fragment g < 4 RB KB is A's line block g / KB, k block g % KB (of the
// step); g >= 4 RB KB is B's line block (g - 4 RB KB) / KB, k block g % KB.
// A wave w < 4 loads A's line blocks RB w + p / KB (p < LOADS_A); a wave
// w >= 4 loads B's line blocks 2 (w % 4) + p / KB (p < LOADS_B); k block p % KB,
// by buffer loads that write lane L's 16 bytes at +16 L.
F4J_HD bool loads_a(int w) { return w < WAVES / 2; }
// Base (bytes into the operand's slice, 64-bit) of wave w's loads at step t for a block
// whose operand lines start at x0 (m0 for A, n0 for B), and the per-load offset.
F4J_HD int64_t load_base(int64_t x0, int w, int t, int KC) {
    const int lb = loads_a(w) ? RB * (w & 3) : 2 * (w & 3);
    return ((((x0 >> 5) + lb) * KC + KB * (int64_t)t) << 10);
}
F4J_HD int load_soff(int p, int KC) { return (((p / KB) * KC) + (p % KB)) << 10; }
F4J_HD int load_voff(int lane) { return 16 * lane; }
F4J_HD int load_lds(int slot, int w, int p) {
    const int g = loads_a(w) ? LOADS_A * w + p : 4 * LOADS_A + LOADS_B * (w & 3) + p;
    return (slot % STAGES) * STAGE + g * FRAG;
}
// Wave w computes rows 32 RB (w % 4) .. + 32 RB - 1 (line blocks i = 0 .. RB - 1) and columns
// 128 (w / 4) .. + 127 (line blocks j = 0..3) of the tile.  Fragment reads at k step ks (0 .. KB - 1):
F4J_HD int a_frag(int slot, int w, int i, int ks, int lane) {
    return (slot % STAGES) * STAGE + ((RB * (w & 3) + i) * KB + ks) * FRAG + 16 * lane;
}
F4J_HD int b_frag(int slot, int w, int j, int ks, int lane) {
    return (slot % STAGES) * STAGE + (4 * RB * KB + (4 * (w >> 2) + j) * KB + ks) * FRAG + 16 * lane;
}
F4J_HD int wave_row(int w) { return 32 * RB * (w & 3); }
F4J_HD int wave_col(int w) { return 128 * (w >> 2); }
// MFMA result item r of `lane` in block (i, j): row and column within the wave tile.
F4J_HD int acc_row(int i, int r, int lane) { return 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
F4J_HD int acc_col(int j, int lane) { return 32 * j + (lane & 31); }

// Product residues of one modulus: a plane of Mp x Np bytes in the MFMA's result
// order, so the epilogue stores a lane's 16 residues of a 32 x 32 block with one
// 16-byte store.  Block (row / 32, col / 32) is 1 KB at res_block, blocks ordered
// with the column block fastest (NB = Np / 32 per block row); lane L's bytes
// 16 L .. 16 L + 15 are its result items r = 0 .. 15, at row res_row(r, L) and
// column res_col(L) of the block (acc_row / acc_col of block (0, 0)).
F4J_HD int64_t res_block(int64_t row0, int64_t col0, int64_t NB) { return ((row0 >> 5) * NB + (col0 >> 5)) << 10; }
F4J_HD int res_row(int r, int lane) { return acc_row(0, r, lane); }
F4J_HD int res_col(int lane) { return acc_col(0, lane); }

}  // namespace crt
}  // namespace f4j
