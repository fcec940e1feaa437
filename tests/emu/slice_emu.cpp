// Host-side emulator of the fragment-major digit-slice product kernel's index math.
//
// Runs the digit split kernels' writes (contiguous lines: 8 digits of one line per
// thread; strided lines: 64 x 64 blocks transposed, 8 digits per item), and the
// product kernel's loads into the LDS ring, fragment reads, FP32 accumulation with
// its flushes into FP64, and epilogue, sequentially on the CPU with the kernel's own
// offset functions (flag4j_amd/csrc/slice_frag.h) and tile mapping (tile_of).  The
// MFMA is emulated from the documented semantics of v_mfma_f32_32x32x16_{bf16,f16},
// written here independently of slice_frag.h: operand lane L holds line L % 32,
// k = 8 (L / 32) + v; result item r of lane L is row 8 (r / 4) + 4 (L / 32) + r % 4,
// column L % 32.  Digits are small integers, so every sum is exact and the result
// is compared with a direct integer sum.  Checks: every digit of the padded slices
// written once and in bounds, every load in bounds, every fragment read from the LDS
// stage of the step being computed, every FP32 partial sum below 2^24 (exact), every
// C entry written once and equal to sum over the digit pairs of 2^shift A_s B_t^T.
//
// usage: slice_emu <M> <N> <K> <S> <flush_steps> [seed]
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../flag4j_amd/csrc/gemm_tile.h"
#include "../../flag4j_amd/csrc/slice_frag.h"

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

// S digit slices of an nx x K operand as the split kernels write them, in
// digit units (byte offset / 2).
static std::vector<int> split_write(const std::vector<std::vector<int>> &X, int64_t nx, int64_t K, int64_t Xp,
                                    int64_t Kp, int S, bool cols) {
    const int64_t KC = 2 * Kp / 32, slice = Xp * Kp;
    std::vector<int> out((size_t)S * slice, 0), hits((size_t)S * slice, 0);
    auto put8 = [&](int64_t x, int64_t k0) {
        for (int s = 0; s < S; ++s) {
            const int64_t b = dfrag::frag_offset(x, 2 * k0, KC);
            CHECK(b % 16 == 0, "split store of x=%ld k=%ld not 16-byte aligned", (long)x, (long)k0);
            for (int u = 0; u < 8; ++u) {
                const int64_t d = (b + 2 * u) / 2;
                CHECK(d >= 0 && d < slice, "split store out of bounds");
                if (d < 0 || d >= slice) continue;
                const int v = (x < nx && k0 + u < K) ? X[s][x * K + k0 + u] : 0;
                out[s * slice + d] = v;
                ++hits[s * slice + d];
            }
        }
    };
    if (!cols) {
        for (int64_t g = 0; g < Xp * (Kp / 8); ++g) put8(g / (Kp / 8), (g % (Kp / 8)) * 8);
    } else {
        for (int64_t xb = 0; xb < Xp / 64; ++xb)
            for (int64_t kb = 0; kb < Kp / 64; ++kb)
                for (int item = 0; item < 512; ++item) put8(xb * 64 + (item >> 3), kb * 64 + (item & 7) * 8);
    }
    for (size_t i = 0; i < hits.size(); ++i) CHECK(hits[i] == 1, "slice digit %zu written %d times", i, hits[i]);
    return out;
}

int main(int argc, char **argv) {
    if (argc < 6) {
        std::printf("usage: slice_emu M N K S flush_steps [seed]\n");
        return 2;
    }
    const int64_t M = std::atol(argv[1]), N = std::atol(argv[2]), K = std::atol(argv[3]);
    const int S = std::atoi(argv[4]), FLUSH = std::atoi(argv[5]);
    const unsigned seed = argc > 6 ? (unsigned)std::atoi(argv[6]) : 1u;
    const int64_t Mp = (M + dfrag::BM - 1) / dfrag::BM * dfrag::BM, Np = (N + dfrag::BN - 1) / dfrag::BN * dfrag::BN;
    const int64_t Kp = (K + 255) / 256 * 256, Kpb = 2 * Kp, KC = Kpb / 32;
    std::mt19937 g(seed);
    std::uniform_int_distribution<int> d(-256, 256);       // BF16 digits lie in [-256, 256]
    std::vector<std::vector<int>> A(S, std::vector<int>(M * K)), B(S, std::vector<int>(N * K));
    for (auto &v : A)
        for (auto &x : v) x = d(g);
    for (auto &v : B)
        for (auto &x : v) x = d(g);
    const std::vector<int> As = split_write(A, M, K, Mp, Kp, S, false), Bs = split_write(B, N, K, Np, Kp, S, true);
    // digit pairs (s, t) of the triangle s + t <= S - 1 (0-based), shift = s + t (a stand-in weight)
    std::vector<int> sa, sb, sh;
    for (int L = S - 1; L >= 0; --L)
        for (int s = 0; s <= L; ++s) {
            sa.push_back(s);
            sb.push_back(L - s);
            sh.push_back(S - 1 - L);
        }
    const int nseg = (int)sa.size();
    const int tps = (int)(Kpb / dfrag::SK), nt = nseg * tps;

    std::vector<double> C(M * N, 0.0);
    std::vector<int> Chits(M * N, 0);
    const int64_t tiles_m = Mp / dfrag::BM, tiles_n = Np / dfrag::BN;
    std::vector<int> lds(dfrag::STAGES * dfrag::STAGE / 2), tag(dfrag::STAGES * dfrag::STAGE / 2);
    for (int64_t bid = 0; bid < tiles_m * tiles_n; ++bid) {
        int64_t tm, tn;
        tile_of(bid, tiles_m, tiles_n, tm, tn);
        const int64_t m0 = tm * dfrag::BM, n0 = tn * dfrag::BN;
        std::fill(tag.begin(), tag.end(), -1);
        auto issue = [&](int slot, int t) {
            const int seg = t / tps;
            const int64_t kt = t % tps;
            for (int w = 0; w < dfrag::WAVES; ++w) {
                const bool is_a = w < dfrag::WAVES / 2;      // restated, not dfrag::loads_a
                const std::vector<int> &src = is_a ? As : Bs;
                const int64_t slice = (is_a ? Mp : Np) * Kp;
                const int64_t base = (is_a ? sa[seg] : sb[seg]) * slice;
                const int64_t x0 = is_a ? m0 : n0;
                for (int p = 0; p < (is_a ? dfrag::LOADS_A : dfrag::LOADS_B); ++p)
                    for (int lane = 0; lane < 64; ++lane)
                        for (int v = 0; v < 8; ++v) {
                            const int64_t byte = dfrag::load_src(x0, w, p, kt, KC) + dfrag::load_voff(lane) + 2 * v;
                            const int dst = (dfrag::load_lds(slot, w, p) + 16 * lane + 2 * v) / 2;
                            CHECK(byte >= 0 && byte / 2 < slice, "%c load out of bounds", is_a ? 'A' : 'B');
                            CHECK(dst >= 0 && dst < (int)lds.size(), "LDS write out of bounds");
                            if (byte < 0 || byte / 2 >= slice || dst < 0 || dst >= (int)lds.size()) continue;
                            lds[dst] = src[base + byte / 2];
                            tag[dst] = t;
                        }
            }
        };
        auto read = [&](int byte_off, int t) -> int {
            const int i = byte_off / 2;
            CHECK(i >= 0 && i < (int)lds.size(), "LDS read out of bounds");
            if (i < 0 || i >= (int)lds.size()) return 0;
            CHECK(tag[i] == t, "LDS read of step %d finds step %d", t, tag[i]);
            return lds[i];
        };
        // acc32[w][i][j][lane][r] (exact integers, checked below 2^24), acc64 likewise as doubles
        const size_t NACC = (size_t)dfrag::WAVES * 2 * 2 * 64 * 16;
        std::vector<long long> acc32(NACC, 0);
        std::vector<double> acc64(NACC, 0.0);
        // per-wave flush counters: waves 4-7 start half an interval in (the kernel's stagger)
        std::vector<int> since(dfrag::WAVES);
        for (int w = 0; w < dfrag::WAVES; ++w) since[w] = w < dfrag::WAVES / 2 ? 0 : FLUSH / 2;
        for (int t = 0; t < dfrag::STAGES - 1; ++t) issue(t, std::min(t, nt - 1));
        for (int t = 0; t < nt; ++t) {
            issue(t + dfrag::STAGES - 1, std::min(t + dfrag::STAGES - 1, nt - 1));   // into the stage step t does not read
            for (int w = 0; w < dfrag::WAVES; ++w)
                for (int ks = 0; ks < dfrag::KB; ++ks) {
                    int a[2][64][8], b[2][64][8];
                    for (int lane = 0; lane < 64; ++lane)
                        for (int v = 0; v < 8; ++v)
                            for (int i = 0; i < 2; ++i) {
                                a[i][lane][v] = read(dfrag::a_frag(t, w, i, ks, lane) + 2 * v, t);
                                b[i][lane][v] = read(dfrag::b_frag(t, w, i, ks, lane) + 2 * v, t);
                            }
                    for (int i = 0; i < 2; ++i)
                        for (int j = 0; j < 2; ++j)
                            for (int lane = 0; lane < 64; ++lane)
                                for (int r = 0; r < 16; ++r) {
                                    const int row = 8 * (r / 4) + 4 * (lane / 32) + r % 4, col = lane % 32;
                                    long long s = 0;
                                    for (int kk = 0; kk < 16; ++kk)
                                        s += (long long)a[i][row + 32 * (kk / 8)][kk % 8] * b[j][col + 32 * (kk / 8)][kk % 8];
                                    long long &x = acc32[((((size_t)w * 2 + i) * 2 + j) * 64 + lane) * 16 + r];
                                    x += s;
                                    CHECK(x < (1LL << 24) && x > -(1LL << 24), "FP32 partial sum beyond 2^24");
                                }
                }
            // each wave flushes every FLUSH steps of its own count and at the end of each digit pair
            for (int w = 0; w < dfrag::WAVES; ++w)
                if (++since[w] == FLUSH || (t + 1) % tps == 0) {
                    const double sc = (double)(1LL << sh[t / tps]);
                    const size_t q0 = (size_t)w * 2 * 2 * 64 * 16, q1 = q0 + 2 * 2 * 64 * 16;
                    for (size_t q = q0; q < q1; ++q) {
                        acc64[q] += (double)acc32[q] * sc;
                        acc32[q] = 0;
                    }
                    since[w] = 0;
                }
        }
        for (int w = 0; w < dfrag::WAVES; ++w)
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                    for (int lane = 0; lane < 64; ++lane)
                        for (int r = 0; r < 16; ++r) {
                            const int64_t row = m0 + dfrag::wave_row(w) + dfrag::acc_row(i, r, lane);
                            const int64_t col = n0 + dfrag::wave_col(w) + dfrag::acc_col(j, lane);
                            CHECK(acc32[((((size_t)w * 2 + i) * 2 + j) * 64 + lane) * 16 + r] == 0, "unflushed sum");
                            if (row >= M || col >= N) continue;
                            C[row * N + col] = acc64[((((size_t)w * 2 + i) * 2 + j) * 64 + lane) * 16 + r];
                            ++Chits[row * N + col];
                        }
    }
    for (int64_t x = 0; x < M; ++x)
        for (int64_t y = 0; y < N; ++y) {
            double want = 0.0;
            for (int q = 0; q < nseg; ++q) {
                long long s = 0;
                for (int64_t k = 0; k < K; ++k) s += (long long)A[sa[q]][x * K + k] * B[sb[q]][y * K + k];
                want += (double)s * (double)(1LL << sh[q]);
            }
            CHECK(Chits[x * N + y] == 1, "C(%ld,%ld) written %d times", (long)x, (long)y, Chits[x * N + y]);
            CHECK(C[x * N + y] == want, "C(%ld,%ld) = %.17g, want %.17g", (long)x, (long)y, C[x * N + y], want);
        }
    if (g_fail) {
        std::printf("FAILED %d checks\n", g_fail);
        return 1;
    }
    std::printf("OK M=%ld N=%ld K=%ld S=%d flush=%d\n", (long)M, (long)N, (long)K, S, FLUSH);
    return 0;
}
