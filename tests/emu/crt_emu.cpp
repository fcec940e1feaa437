// Host-side emulator of the modular (CRT) GEMM's index math.
//
// Runs the residue split kernel's writes and the INT8 product kernel's loads,
// LDS stages, fragment reads and epilogue sequentially on the CPU, using the
// kernels' own offset functions (flag4j_amd/csrc/crt_layout.h) and tile
// mapping (tile_of).  The MFMA is emulated from the documented semantics of
// v_mfma_i32_32x32x32_i8, written here independently of crt_layout.h: operand
// lane L holds line L % 32, k = 16 (L / 32) + v; result item r of lane L is
// row 8 (r / 4) + 4 (L / 32) + r % 4, column L % 32.  Checks: every residue
// byte written once and in bounds, every load in bounds, every fragment read
// from the LDS stage of the k slice being computed, every product residue
// written once and equal to the direct integer product modulo m.
//
// usage: crt_emu <M> <N> <K> [seed]
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../flag4j_amd/csrc/gemm_tile.h"
#include "../../flag4j_amd/csrc/crt_layout.h"

using namespace f4j;

static int g_fail = 0;
#define CHECK(c, ...)                              \
    do {                                           \
        if (!(c)) {                                \
            if (g_fail++ < 10) {                   \
                std::printf("FAIL: " __VA_ARGS__); \
                std::printf("\n");                 \
            }                                      \
        }                                          \
    } while (0)

// Residue slice as the split kernels write it: contiguous lines in blocks of
// 32 lines x 128 k, strided lines (cols) in blocks of 128 lines x 32 k.
static std::vector<signed char> split_write(const std::vector<int> &X, int64_t nx, int64_t K, int64_t Xp,
                                            int64_t Kp, bool cols) {
    const int64_t KC = Kp / 32;
    const int bx = cols ? 128 : 32, bk = cols ? 32 : 128;
    std::vector<signed char> out(Xp * Kp, 0);
    std::vector<int> hits(Xp * Kp, 0);
    for (int64_t xb = 0; xb < Xp / bx; ++xb)
        for (int64_t kb = 0; kb < Kp / bk; ++kb)
            for (int t = 0; t < 256; ++t) {
                const int64_t x = xb * bx + (cols ? crt::split_cols_line(t) : crt::split_line(t));
                const int64_t k = kb * bk + (cols ? crt::split_cols_k(t) : crt::split_k(t));
                const int64_t off = crt::frag_offset(x, k, KC);
                for (int u = 0; u < 16; ++u) {
                    CHECK(off + u >= 0 && off + u < Xp * Kp, "split write out of bounds x=%ld k=%ld", (long)x,
                          (long)(k + u));
                    if (off + u < 0 || off + u >= Xp * Kp) continue;
                    const int v = (x < nx && k + u < K) ? X[x * K + k + u] : 0;
                    out[off + u] = (signed char)v;
                    ++hits[off + u];
                }
            }
    for (int64_t i = 0; i < Xp * Kp; ++i) CHECK(hits[i] == 1, "residue byte %ld written %d times", (long)i, hits[i]);
    return out;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        std::printf("usage: crt_emu M N K [seed]\n");
        return 2;
    }
    const int64_t M = std::atol(argv[1]), N = std::atol(argv[2]), K = std::atol(argv[3]);
    const unsigned seed = argc > 4 ? (unsigned)std::atoi(argv[4]) : 1u;
    const int m = 251;
    const int64_t Mp = (M + crt::BM - 1) / crt::BM * crt::BM, Np = (N + crt::BN - 1) / crt::BN * crt::BN;
    const int64_t Kp = (K + crt::BK - 1) / crt::BK * crt::BK;
    const int KC = (int)(Kp / 32), nt = (int)(Kp / crt::BK);

    std::mt19937 g(seed);
    std::uniform_int_distribution<int> d(-128, 127);
    std::vector<int> A(M * K), B(N * K);            // A: M lines of K; B: N lines (columns of op(B)) of K
    for (auto &v : A) v = d(g);
    for (auto &v : B) v = d(g);
    // A through the contiguous-line split, B through the strided one (as for op(B) = N)
    const std::vector<signed char> Ar = split_write(A, M, K, Mp, Kp, false), Br = split_write(B, N, K, Np, Kp, true);

    std::vector<int> R(M * N, -1), Rhits(M * N, 0);
    const int64_t tiles_m = Mp / crt::BM, tiles_n = Np / crt::BN;
    std::vector<int> tile_hits(tiles_m * tiles_n, 0);
    std::vector<signed char> lds(crt::STAGES * crt::STAGE);
    std::vector<int> tag(crt::STAGES * crt::STAGE, -1);

    for (int64_t bid = 0; bid < tiles_m * tiles_n; ++bid) {
        int64_t tm, tn;
        tile_of(bid, tiles_m, tiles_n, tm, tn);
        CHECK(tm >= 0 && tm < tiles_m && tn >= 0 && tn < tiles_n, "tile out of range");
        ++tile_hits[tm * tiles_n + tn];
        const int64_t n0 = tn * crt::BN;
        std::fill(tag.begin(), tag.end(), -1);
        auto issue_b = [&](int slot, int t) {
            for (int w = 0; w < 4; ++w)
                for (int p = 0; p < 4; ++p)
                    for (int lane = 0; lane < 64; ++lane)
                        for (int v = 0; v < 16; ++v) {
                            const int64_t src = crt::b_base(n0, t, KC) + crt::b_soff(w, p, KC) + crt::b_voff(lane) + v;
                            const int dst = crt::b_lds(slot, w, p) + 16 * lane + v;
                            CHECK(src >= 0 && src < Np * Kp, "B load out of bounds");
                            CHECK(dst >= 0 && dst < crt::STAGES * crt::STAGE, "LDS write out of bounds");
                            if (src < 0 || src >= Np * Kp || dst < 0 || dst >= crt::STAGES * crt::STAGE) continue;
                            lds[dst] = Br[src];
                            tag[dst] = t;
                        }
        };
        // acc[w][i][j][lane][r]
        std::vector<long long> acc(4 * 2 * 4 * 64 * 16, 0);
        auto A_at = [&](int w, int i, int t, int ks, int lane, int v) -> int {
            const int64_t m0w = tm * crt::BM + 64 * w;
            const int64_t off = crt::a_base(m0w, t, KC) + crt::a_soff(i, KC) + crt::a_voff(ks, lane) + v;
            CHECK(off >= 0 && off < Mp * Kp, "A load out of bounds");
            return (off >= 0 && off < Mp * Kp) ? Ar[off] : 0;
        };
        auto B_at = [&](int t, int j, int ks, int lane, int v) -> int {
            const int off = crt::b_frag(t, j, ks, lane) + v;
            CHECK(off >= 0 && off < crt::STAGES * crt::STAGE, "LDS read out of bounds");
            if (off < 0 || off >= crt::STAGES * crt::STAGE) return 0;
            CHECK(tag[off] == t, "LDS read of slice %d finds slice %d", t, tag[off]);
            return lds[off];
        };
        issue_b(0, 0);
        issue_b(1, std::min(1, nt - 1));
        for (int t = 0; t < nt; ++t) {
            issue_b(t + 2, std::min(t + 2, nt - 1));    // into the stage no wave reads in step t
            for (int w = 0; w < 4; ++w)
                for (int ks = 0; ks < 4; ++ks) {
                    // operands as the lanes hold them
                    int a[2][64][16], b[4][64][16];
                    for (int lane = 0; lane < 64; ++lane)
                        for (int v = 0; v < 16; ++v) {
                            for (int i = 0; i < 2; ++i) a[i][lane][v] = A_at(w, i, t, ks, lane, v);
                            for (int j = 0; j < 4; ++j) b[j][lane][v] = B_at(t, j, ks, lane, v);
                        }
                    // documented MFMA semantics
                    for (int i = 0; i < 2; ++i)
                        for (int j = 0; j < 4; ++j)
                            for (int lane = 0; lane < 64; ++lane)
                                for (int r = 0; r < 16; ++r) {
                                    const int row = 8 * (r / 4) + 4 * (lane / 32) + r % 4, col = lane % 32;
                                    long long s = 0;
                                    for (int kk = 0; kk < 32; ++kk) {
                                        const int la = row + 32 * (kk / 16), lb = col + 32 * (kk / 16);
                                        s += (long long)a[i][la][kk % 16] * b[j][lb][kk % 16];
                                    }
                                    acc[(((w * 2 + i) * 4 + j) * 64 + lane) * 16 + r] += s;
                                }
                }
        }
        // epilogue
        for (int w = 0; w < 4; ++w)
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 4; ++j)
                    for (int lane = 0; lane < 64; ++lane)
                        for (int r = 0; r < 16; ++r) {
                            const int64_t row = tm * crt::BM + 64 * w + crt::acc_row(i, r, lane);
                            const int64_t col = n0 + crt::acc_col(j, lane);
                            if (row >= M || col >= N) continue;
                            long long v = acc[(((w * 2 + i) * 4 + j) * 64 + lane) * 16 + r] % m;
                            if (v < 0) v += m;
                            R[row * N + col] = (int)v;
                            ++Rhits[row * N + col];
                        }
    }
    for (size_t i = 0; i < tile_hits.size(); ++i) CHECK(tile_hits[i] == 1, "tile %zu visited %d times", i, tile_hits[i]);
    for (int64_t x = 0; x < M; ++x)
        for (int64_t y = 0; y < N; ++y) {
            long long s = 0;
            for (int64_t k = 0; k < K; ++k) s += (long long)A[x * K + k] * B[y * K + k];
            s %= m;
            if (s < 0) s += m;
            CHECK(Rhits[x * N + y] == 1, "R(%ld,%ld) written %d times", (long)x, (long)y, Rhits[x * N + y]);
            CHECK(R[x * N + y] == s, "R(%ld,%ld) = %d, want %lld", (long)x, (long)y, R[x * N + y], s);
        }
    if (g_fail) {
        std::printf("FAILED %d checks\n", g_fail);
        return 1;
    }
    std::printf("OK M=%ld N=%ld K=%ld\n", (long)M, (long)N, (long)K);
    return 0;
}
