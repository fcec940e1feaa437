// Index math of the blocked matrix-vector kernels (gemv_kernel, gemv_dma_kernel
// in gemv.hip), shared with the host emulator that runs them on the CPU
// (tests/emu/gemv_emu.cpp).
//
// A wave runs 64 chains, chain c = (row c / 8, 64-wide k block c % 8) of an
// 8-row slab.  Stage s covers k = kstage(s) + [0, 512): 16 consecutive
// elements of each of the 8 blocks, so four stages finish a 512-wide k window
// and every chain's block sum.  The elements of a stage reach LDS in the order
// the loads deliver them (lane-major), rotated per lane so that the 64 readers
// (one chain each) spread over the banks.
#pragma once
#include <cstdint>

#ifndef F4J_HD
#define F4J_HD __host__ __device__ inline
#endif

namespace f4j {
namespace gemv {

constexpr int ROWS = 8;                   // rows per wave (8 rows x 8 blocks = 64 chains)
constexpr int SPAN = 64 * 7 + 16;         // a stage is whole when kstage(s) + SPAN <= K

F4J_HD int kstage(int64_t s) { return (int)(512 * (s >> 2) + 16 * (s & 3)); }
F4J_HD int64_t stages_for(int64_t K) { return 4 * ((K + 511) / 512); }

// Row (within the slab) and k offset (within the stage) of A load i of lane l.
// VEC 2: 16-byte loads of pairs (8 loads per stage), VEC 1: 8-byte loads (16).
// The reader of chain c finds pair p at (p + c/8 + c%8) & 7 (VEC 2), element e at
// (e + c/4 + c%4) & 15 (VEC 1) of its 16-element run.
template <int VEC>
F4J_HD void a_slot(int i, int l, int &r, int &kl) {
    int ch, el;
    if (VEC == 2) {
        const int b = l >> 3;
        ch = 8 * i + b;
        el = 2 * (((l & 7) - i - b) & 7);
    } else {
        const int b = l >> 4;
        ch = 4 * i + b;
        el = ((l & 15) - i - b) & 15;
    }
    r = ch >> 3;
    kl = 64 * (ch & 7) + el;
}

// ---- register-staged form (gemv_kernel): LDS indices in doubles ----

// k offset (within the stage) of x load h (of 2) of lane l: block j = 4 h + l / 16,
// element e of the block's run at lane position (e + j) & 15.
F4J_HD int x_slot(int h, int l) {
    const int j = 4 * h + (l >> 4);
    return 64 * j + (((l & 15) - j) & 15);
}
// Where load i (A) / h (x) of lane l is written.
template <int VEC>
F4J_HD int a_store(int i, int l) { return VEC == 2 ? 128 * i + 2 * l : 64 * i + l; }
F4J_HD int x_store(int h, int l) { return 1024 + 64 * h + l; }
// Where chain l reads: A element e (VEC 1) or pair p (VEC 2), and x element e.
template <int VEC>
F4J_HD int a_read(int l, int e) {
    return VEC == 2 ? 128 * (l >> 3) + 16 * (l & 7) + 2 * ((e + (l >> 3) + (l & 7)) & 7)
                    : 64 * (l >> 2) + 16 * (l & 3) + ((e + (l >> 2) + (l & 3)) & 15);
}
F4J_HD int x_read(int l, int e) {
    const int j = l & 7;
    return 1024 + 64 * (j >> 2) + 16 * (j & 3) + ((e + j) & 15);
}

// ---- direct-to-LDS form (gemv_dma_kernel): LDS offsets in bytes within a stage slot ----

constexpr int DMA_NL = 8;                                  // 16-byte A loads per lane and stage
constexpr int slot_bytes() { return 1024 * DMA_NL + 1024; }
constexpr int XOFF = 1024 * DMA_NL;                        // x area of a slot
// x loads per stage: XV 2 = 16-byte pairs (x aligned, unit stride, K even), XV 1 = dwords.
template <int XV>
constexpr int x_dma_loads() { return XV == 1 ? 4 : 1; }
template <int XV>
constexpr int dma_loads() { return DMA_NL + x_dma_loads<XV>(); }
// vmcnt at the top of step s of a D-slot ring: the loads of stages s + 1 .. s + D - 2
// may stay in flight, stage s's are retired.
constexpr int dma_wait(int D, int loads) { return (D - 2) * loads; }

// x load h of lane l: k offset within the stage and byte offset within the element
// (XV 1: each dword load moves half of an x element).
template <int XV>
F4J_HD void dma_x_slot(int h, int l, int &xkl, int &xadd) {
    if (XV == 2) {
        const int j = l >> 3;                              // block j: pair p at lane position (p + j) & 7
        xkl = 64 * j + 2 * (((l & 7) - j) & 7);
        xadd = 0;
    } else {
        const int m = 32 * h + (l >> 1);                   // double m of the stage's x area
        const int j = m >> 4;                              // block j: element e at (e + j) & 15
        xkl = 64 * j + (((m & 15) - j) & 15);
        xadd = 4 * (l & 1);
    }
}
// LDS destination of a load: lane l of load i (A, 16 bytes each) / h (x).
F4J_HD int dma_a_store(int i, int l) { return 1024 * i + 16 * l; }
template <int XV>
F4J_HD int dma_x_store(int h, int l) { return XV == 2 ? XOFF + 16 * l : XOFF + 256 * h + 4 * l; }
// Where chain l reads pair p (k = kstage + 64 (l % 8) + 2 p, 2 p + 1): A bytes, x bytes.
F4J_HD int dma_a_read(int l, int p) {
    const int j = l & 7;
    return 1024 * (l >> 3) + 128 * j + 16 * ((p + (l >> 3) + j) & 7);
}
template <int XV>
F4J_HD int dma_x_read(int l, int p, int half) {
    const int j = l & 7;
    return XV == 2 ? XOFF + 128 * j + 16 * ((p + j) & 7) + 8 * half
                   : XOFF + 128 * j + 8 * ((2 * p + half + j) & 15);
}

}  // namespace gemv
}  // namespace f4j
