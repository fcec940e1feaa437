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
// from the LDS stage of the k slice being computed, every product residue byte
// of the padded plane written once, and every entry, read back as the
// reconstruction kernel reads the plane, equal to the direct integer product
// modulo m.
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
    const int KC = (int)(Kp / 32);

    std::mt19937 g(seed);
    std::uniform_int_distribution<int> d(-128, 127);
    std::vector<int> A(M * K), B(N * K);            // A: M lines of K; B: N lines (columns of op(B)) of K
    for (auto &v : A) v = d(g);
    for (auto &v : B) v = d(g);
    // A through the contiguous-line split, B through the strided one (as for op(B) = N)
    const std::vector<signed char> Ar = split_write(A, M, K, Mp, Kp, false), Br = split_write(B, N, K, Np, Kp, true);

    // product residues of one modulus, in the padded plane layout (crt::res_block)
    const int64_t NB = Np / 32;
    std::vector<int> R(Mp * Np, -1), Rhits(Mp * Np, 0);
    const int64_t tiles_m = Mp / crt::BM, tiles_n = Np / crt::BN;
    const int nt = (int)(Kp / crt::SK);
    std::vector<int> tile_hits(tiles_m * tiles_n, 0);
    std::vector<signed char> lds(crt::STAGES * crt::STAGE);
    std::vector<int> tag(crt::STAGES * crt::STAGE, -1);

    for (int64_t bid = 0; bid < tiles_m * tiles_n; ++bid) {
        int64_t tm, tn;
        tile_of(bid, tiles_m, tiles_n, tm, tn);
        CHECK(tm >= 0 && tm < tiles_m && tn >= 0 && tn < tiles_n, "tile out of range");
        ++tile_hits[tm * tiles_n + tn];
        const int64_t m0 = tm * crt::BM, n0 = tn * crt::BN;
        std::fill(tag.begin(), tag.end(), -1);
        // every wave's loads of step t into stage slot (waves 0-3 from A, 4-7 from B)
        auto issue = [&](int slot, int t) {
            for (int w = 0; w < crt::WAVES; ++w) {
                const bool is_a = w < crt::WAVES / 2;          // restated, not crt::loads_a
                const std::vector<signed char> &src = is_a ? Ar : Br;
                const int64_t x0 = is_a ? m0 : n0, size = is_a ? Mp * Kp : Np * Kp;
                for (int p = 0; p < (is_a ? crt::LOADS_A : crt::LOADS_B); ++p)
                    for (int lane = 0; lane < 64; ++lane)
                        for (int v = 0; v < 16; ++v) {
                            const int64_t off = crt::load_base(x0, w, t, KC) + crt::load_soff(p, KC) +
                                                crt::load_voff(lane) + v;
                            const int dst = crt::load_lds(slot, w, p) + 16 * lane + v;
                            CHECK(off >= 0 && off < size, "%c load out of bounds", is_a ? 'A' : 'B');
                            CHECK(dst >= 0 && dst < crt::STAGES * crt::STAGE, "LDS write out of bounds");
                            if (off < 0 || off >= size || dst < 0 || dst >= crt::STAGES * crt::STAGE) continue;
                            lds[dst] = src[off];
                            tag[dst] = t;
                        }
            }
        };
        auto read = [&](int off, int t) -> int {
            CHECK(off >= 0 && off < crt::STAGES * crt::STAGE, "LDS read out of bounds");
            if (off < 0 || off >= crt::STAGES * crt::STAGE) return 0;
            CHECK(tag[off] == t, "LDS read of step %d finds step %d", t, tag[off]);
            return lds[off];
        };
        // acc[w][i][j][lane][r]
        std::vector<long long> acc(crt::WAVES * crt::RB * 4 * 64 * 16, 0);
        for (int t = 0; t < crt::AHEAD; ++t) issue(t, std::min(t, nt - 1));
        for (int t = 0; t < nt; ++t) {
            issue(t + crt::AHEAD, std::min(t + crt::AHEAD, nt - 1));   // into the stage no wave reads in step t
            for (int w = 0; w < crt::WAVES; ++w)
                for (int ks = 0; ks < crt::KB; ++ks) {
                    // operands as the lanes hold them
                    int a[crt::RB][64][16], b[4][64][16];
                    for (int lane = 0; lane < 64; ++lane)
                        for (int v = 0; v < 16; ++v) {
                            for (int i = 0; i < crt::RB; ++i) a[i][lane][v] = read(crt::a_frag(t, w, i, ks, lane) + v, t);
                            for (int j = 0; j < 4; ++j) b[j][lane][v] = read(crt::b_frag(t, w, j, ks, lane) + v, t);
                        }
                    // documented MFMA semantics
                    for (int i = 0; i < crt::RB; ++i)
                        for (int j = 0; j < 4; ++j)
                            for (int lane = 0; lane < 64; ++lane)
                                for (int r = 0; r < 16; ++r) {
                                    const int row = 8 * (r / 4) + 4 * (lane / 32) + r % 4, col = lane % 32;
                                    long long sum = 0;
                                    for (int kk = 0; kk < 32; ++kk) {
                                        const int la = row + 32 * (kk / 16), lb = col + 32 * (kk / 16);
                                        sum += (long long)a[i][la][kk % 16] * b[j][lb][kk % 16];
                                    }
                                    acc[(((w * crt::RB + i) * 4 + j) * 64 + lane) * 16 + r] += sum;
                                }
                }
        }
        // epilogue: lane's 16 residues of block (i, j) as one 16-byte store, byte r = item r
        for (int w = 0; w < crt::WAVES; ++w)
            for (int i = 0; i < crt::RB; ++i)
                for (int j = 0; j < 4; ++j)
                    for (int lane = 0; lane < 64; ++lane) {
                        const int64_t base = crt::res_block(m0 + crt::wave_row(w) + 32 * i,
                                                            n0 + crt::wave_col(w) + 32 * j, NB) + 16 * lane;
                        CHECK(base % 16 == 0 && base >= 0 && base + 16 <= Mp * Np, "residue store out of bounds");
                        if (base < 0 || base + 16 > Mp * Np) continue;
                        for (int r = 0; r < 16; ++r) {
                            long long v = acc[(((w * crt::RB + i) * 4 + j) * 64 + lane) * 16 + r] % m;
                            if (v < 0) v += m;
                            R[base + r] = (int)v;
                            ++Rhits[base + r];
                        }
                    }
    }
    for (int64_t i = 0; i < Mp * Np; ++i) CHECK(Rhits[i] == 1, "residue byte %ld written %d times", (long)i, Rhits[i]);
    // the reconstruction's reading: block b of the plane, lane L, item r -> entry
    // (32 (b / NB) + res_row(r, L), 32 (b % NB) + res_col(L)), checked against the
    // documented MFMA result order (item r of lane L: row 8 (r / 4) + 4 (L / 32) + r % 4, column L % 32)
    std::vector<int> Rd(M * N, -1), Rdh(M * N, 0);
    for (int64_t b = 0; b < (Mp / 32) * NB; ++b)
        for (int lane = 0; lane < 64; ++lane)
            for (int r = 0; r < 16; ++r) {
                CHECK(crt::res_row(r, lane) == 8 * (r / 4) + 4 * (lane / 32) + r % 4 && crt::res_col(lane) == lane % 32,
                      "reconstruction reads item %d of lane %d at the wrong entry", r, lane);
                const int64_t row = 32 * (b / NB) + crt::res_row(r, lane), col = 32 * (b % NB) + crt::res_col(lane);
                if (row >= M || col >= N) continue;
                Rd[row * N + col] = R[(b << 10) + 16 * lane + r];
                ++Rdh[row * N + col];
            }
    for (size_t i = 0; i < tile_hits.size(); ++i) CHECK(tile_hits[i] == 1, "tile %zu visited %d times", i, tile_hits[i]);
    for (int64_t x = 0; x < M; ++x)
        for (int64_t y = 0; y < N; ++y) {
            long long s = 0;
            for (int64_t k = 0; k < K; ++k) s += (long long)A[x * K + k] * B[y * K + k];
            s %= m;
            if (s < 0) s += m;
            CHECK(Rdh[x * N + y] == 1, "R(%ld,%ld) read %d times", (long)x, (long)y, Rdh[x * N + y]);
            CHECK(Rd[x * N + y] == s, "R(%ld,%ld) = %d, want %lld", (long)x, (long)y, Rd[x * N + y], s);
        }
    if (g_fail) {
        std::printf("FAILED %d checks\n", g_fail);
        return 1;
    }
    std::printf("OK M=%ld N=%ld K=%ld\n", (long)M, (long)N, (long)K);
    return 0;
}
