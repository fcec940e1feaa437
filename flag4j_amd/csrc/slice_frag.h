// Index math of the fragment-major digit-slice product kernel (BF16 / FP16
// digits), shared by the kernel and by the host emulator that checks it
// (tests/emu/slice_emu.cpp).
//
// Digit slices are stored as the modular form stores its residues
// (crt_layout.h) with k counted in bytes, two per digit: for slice s of a
// padded Xp x Kp operand (Kpb = 2 Kp bytes a line) the 32 lines x 32 bytes block
// (x / 32, kb / 32) is 1 KB, blocks k-fastest (KC = Kpb / 32 a line block);
// inside, lane L = x % 32 + 32 ((kb / 16) % 2) owns bytes 16 L .. 16 L + 15.
// That is the operand layout of v_mfma_f32_32x32x16_{bf16,f16} (lane holds
// line lane % 32, digits k = 8 (lane / 32) + v), so one 16-byte load per lane
// reads a whole fragment from 1 KB of contiguous memory, and a split thread's 8
// consecutive digits of one line (16 bytes, k a multiple of 8) are one lane's.
//
// Product kernel.  A block computes a BM x BN = 256 x 128 tile with 8 waves of
// 64 x 64 (wave w: rows 64 (w % 4), columns 64 (w / 4); 2 x 2 MFMA blocks, FP32
// accumulators flushed into FP64 ones).  Each step t (k bytes SK t .. SK t + SK - 1
// of one digit pair) both operands' fragments go straight from the slices into
// LDS stage t % STAGES, each fragment once per block: waves 0-3 load A's line
// blocks 2 w, 2 w + 1, waves 4-7 B's line block w - 4, all KB k blocks.
#pragma once
#include <cstdint>

#ifndef F4J_HD
#define F4J_HD __host__ __device__ inline
#endif

namespace f4j {
namespace dfrag {

constexpr int BM = 256, BN = 128;             // block tile (rows, columns)
constexpr int WAVES = 8;
constexpr int SK = 128;                       // k bytes (64 digits) per step
constexpr int KB = SK / 32;                   // 32-byte MFMA k steps (k blocks) per step
constexpr int A_LB = BM / 32, B_LB = BN / 32; // line blocks of A and B in a tile
constexpr int STAGES = 3;
constexpr int LOADS_A = 2 * KB;               // fragment loads per step of an A wave (waves 0-3)
constexpr int LOADS_B = KB;                   // and of a B wave (waves 4-7)
constexpr int STAGE = (A_LB + B_LB) * KB * 1024;   // LDS bytes of one step: A's fragments, then B's

// Byte offset of digit-byte kb of line x in one slice (KC = Kpb / 32).
F4J_HD int64_t frag_offset(int64_t x, int64_t kb, int64_t KC) {
    return (((x >> 5) * KC + (kb >> 5)) << 10) + (((x & 31) + ((kb >> 4) & 1) * 32) << 4) + (kb & 15);
}

F4J_HD bool loads_a(int w) { return w < WAVES / 2; }
// First line block (of the tile's operand lines) that wave w loads.
F4J_HD int load_lb(int w) { return loads_a(w) ? 2 * (w & 3) : (w & 3); }
// Byte offset in the slice of load p (p < LOADS_A or LOADS_B) of wave w at k step kt of
// a digit pair, for a block whose operand lines start at x0.
F4J_HD int64_t load_src(int64_t x0, int w, int p, int64_t kt, int64_t KC) {
    return (((x0 >> 5) + load_lb(w) + p / KB) * KC + KB * kt + p % KB) << 10;
}
F4J_HD int load_voff(int lane) { return 16 * lane; }
// LDS byte offset of load p of wave w into stage `slot`.
F4J_HD int load_lds(int slot, int w, int p) {
    const int g = loads_a(w) ? (load_lb(w) + p / KB) * KB + p % KB : (A_LB + load_lb(w)) * KB + p % KB;
    return (slot % STAGES) * STAGE + g * 1024;
}
// Fragment reads of wave w at k step ks of stage `slot`: A line block 2 (w % 4) + i,
// B line block 2 (w / 4) + j.
F4J_HD int a_frag(int slot, int w, int i, int ks, int lane) {
    return (slot % STAGES) * STAGE + ((2 * (w & 3) + i) * KB + ks) * 1024 + 16 * lane;
}
F4J_HD int b_frag(int slot, int w, int j, int ks, int lane) {
    return (slot % STAGES) * STAGE + ((A_LB + 2 * (w >> 2) + j) * KB + ks) * 1024 + 16 * lane;
}
F4J_HD int wave_row(int w) { return 64 * (w & 3); }
F4J_HD int wave_col(int w) { return 64 * (w >> 2); }
// MFMA result item r of `lane` in block (i, j): row and column within the wave tile.
F4J_HD int acc_row(int i, int r, int lane) { return 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
F4J_HD int acc_col(int j, int lane) { return 32 * j + (lane & 31); }

}  // namespace dfrag
}  // namespace f4j
