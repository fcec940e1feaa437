// Host-side emulator of the FP64 MFMA GEMM kernel's index math.
//
// Runs the same tile mapping (tile_of), stored-block geometry, LDS placement
// and per-lane fragment addressing (Opnd<> from gemm_tile.h) as the device
// kernel, sequentially on the CPU, with the documented operand/accumulator
// layouts of v_mfma_f64_16x16x4f64.  Every LDS write and read is bounds
// checked, every read must hit a location written in the same stage, and
// every C element must be written exactly once.  Inputs are small integers,
// so the product is exact and compared bit for bit with a direct product.
//
// usage: gemm_emu <ka> <kb> <M> <N> <K> <cfg> [vec]   (M,N,K in element units;
//        kinds 0..4 = R_N R_T Z_N Z_T Z_C; vec = 1 or 2 doubles per load)
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../flag4j_amd/csrc/gemm_tile.h"

using namespace f4j;

static int g_fail = 0;
#define CHECK(c, ...)                          \
    do {                                       \
        if (!(c)) {                            \
            if (g_fail++ < 10) {               \
                std::printf("FAIL: " __VA_ARGS__); \
                std::printf("\n");             \
            }                                  \
        }                                      \
    } while (0)

// Buffer-load semantics of TileCopy: a lane's V doubles come from
// base + offset unless offset >= num_records, which reads as 0.  Each value
// must equal the bounds-checked stored element it stands for.
template <class O, int NT, int V>
void copy_tile(const double *G, int64_t ld, int64_t x0, int64_t k0, int64_t X, int64_t K, std::vector<double> &lds,
               std::vector<int> &w) {
    using TC = TileCopy<O, NT, V>;
    int64_t r0, c0, rows, cols;
    O::block(x0, k0, X, K, r0, c0, rows, cols);
    CHECK(r0 < rows && c0 < cols, "block origin outside the matrix");
    const uint32_t nr = TC::num_records(r0, c0, rows, cols, ld);
    const double *base = G + r0 * ld + c0;
    for (int tid = 0; tid < NT; ++tid)
        for (int i = 0; i < TC::PASSES; ++i) {
            const uint32_t off = TC::offset(tid, i, c0, cols, ld);
            CHECK(off == TC::OOB || off % (8 * V) == 0 || V == 1, "misaligned vector offset");
            for (int v = 0; v < V; ++v) {
                const uint64_t o = (uint64_t)off + 8 * v;
                const double got = o + 8 <= nr ? base[o / 8] : 0.0;
                const int64_t r = r0 + TC::lane_row(tid) + i * TC::ROWS_PER_PASS, c = c0 + TC::lane_col(tid) + v;
                const double want = (r < rows && c < cols) ? G[r * ld + c] : 0.0;
                CHECK(got == want, "buffer load (%ld,%ld) got %g want %g", (long)r, (long)c, got, want);
                const int li = TC::lds_index(tid, i) + v;
                CHECK(li >= 0 && li < O::LDS_DOUBLES, "lds write oob");
                if (g_fail) return;
                lds[li] = got;
                w[li]++;
            }
        }
}

template <int BM, int BN, int BK, int WGM, int WGN, int V, int KA, int KB>
void emulate(const GemmProblem &p, std::vector<int> &written) {
    constexpr int NT = WGM * WGN * 64;
    constexpr int WTM = BM / WGM, WTN = BN / WGN;
    constexpr int FM = WTM / 16, FN = WTN / 16;
    using OA = Opnd<KA, true, BM, BK>;
    using OB = Opnd<KB, false, BN, BK>;
    const int64_t tiles_m = (p.M + BM - 1) / BM, tiles_n = (p.N + BN - 1) / BN;
    std::vector<int> tile_seen(tiles_m * tiles_n, 0);
    for (int64_t bid = 0; bid < tiles_m * tiles_n; ++bid) {
        int64_t tm, tn;
        tile_of(bid, tiles_m, tiles_n, tm, tn);
        CHECK(tm >= 0 && tm < tiles_m && tn >= 0 && tn < tiles_n, "tile out of range bid=%ld", (long)bid);
        if (g_fail) return;
        tile_seen[tm * tiles_n + tn]++;
        const int64_t m0 = tm * BM, n0 = tn * BN;
        std::vector<double> acc((size_t)WGM * WGN * FM * FN * 64 * 4, 0.0);
        const int64_t nk = (p.K + BK - 1) / BK;
        for (int64_t kt = 0; kt < nk; ++kt) {
            const int64_t k0 = kt * BK;
            std::vector<double> la(OA::LDS_DOUBLES, 0.0), lb(OB::LDS_DOUBLES, 0.0);
            std::vector<int> wa(OA::LDS_DOUBLES, 0), wb(OB::LDS_DOUBLES, 0);
            copy_tile<OA, NT, V>(p.A, p.lda, m0, k0, p.M, p.K, la, wa);
            copy_tile<OB, NT, V>(p.B, p.ldb, n0, k0, p.N, p.K, lb, wb);
            if (g_fail) return;
            for (int li = 0; li < OA::LDS_DOUBLES; ++li) CHECK(wa[li] <= 1, "A lds double write");
            for (int li = 0; li < OB::LDS_DOUBLES; ++li) CHECK(wb[li] <= 1, "B lds double write");
            for (int wave = 0; wave < WGM * WGN; ++wave) {
                const int wm = (wave / WGN) * WTM, wn = (wave % WGN) * WTN;
                for (int ks = 0; ks < BK / 4; ++ks) {
                    double a[FM][64], b[FN][64];
                    for (int lane = 0; lane < 64; ++lane) {
                        for (int i = 0; i < FM; ++i) {
                            const int li = OA::lane_base(lane, wm) + ks * OA::KS + i * OA::FS;
                            CHECK(li >= 0 && li < OA::LDS_DOUBLES && wa[li] == 1, "A frag read oob/unwritten");
                            if (g_fail) return;
                            a[i][lane] = la[li] * OA::lane_sign(lane);
                        }
                        for (int j = 0; j < FN; ++j) {
                            const int li = OB::lane_base(lane, wn) + ks * OB::KS + j * OB::FS;
                            CHECK(li >= 0 && li < OB::LDS_DOUBLES && wb[li] == 1, "B frag read oob/unwritten");
                            if (g_fail) return;
                            b[j][lane] = lb[li] * OB::lane_sign(lane);
                        }
                    }
                    // v_mfma_f64_16x16x4f64: A[i][k] in lane i+16k, B[k][j] in lane j+16k,
                    // D[i][j] in lane j+16*(i%4), item i/4.
                    for (int i = 0; i < FM; ++i)
                        for (int j = 0; j < FN; ++j)
                            for (int di = 0; di < 16; ++di)
                                for (int dj = 0; dj < 16; ++dj) {
                                    double s = 0;
                                    for (int k = 0; k < 4; ++k) s += a[i][di + 16 * k] * b[j][dj + 16 * k];
                                    const int lane = dj + 16 * (di % 4), item = di / 4;
                                    acc[((((size_t)wave * FM + i) * FN + j) * 64 + lane) * 4 + item] += s;
                                }
                }
            }
        }
        for (int wave = 0; wave < WGM * WGN; ++wave) {
            const int wm = (wave / WGN) * WTM, wn = (wave % WGN) * WTN;
            for (int lane = 0; lane < 64; ++lane)
                for (int i = 0; i < FM; ++i)
                    for (int j = 0; j < FN; ++j) {
                        const int64_t col = n0 + wn + (lane & 15) + j * 16;
                        if (col >= p.N) continue;
                        for (int r = 0; r < 4; ++r) {
                            const int64_t row = m0 + wm + i * 16 + mfma_f64_acc_row(lane, r);
                            if (row < p.M) {
                                p.C[row * p.ldc + col] = acc[((((size_t)wave * FM + i) * FN + j) * 64 + lane) * 4 + r];
                                written[row * p.N + col]++;
                            }
                        }
                    }
        }
    }
    for (size_t t = 0; t < tile_seen.size(); ++t) CHECK(tile_seen[t] == 1, "tile %zu seen %d times", t, tile_seen[t]);
}

template <int V, int KA, int KB>
void run_v(int cfg, const GemmProblem &p, std::vector<int> &w) {
    switch (cfg) {
#define F4J_TILE_CASE(id, BM, BN, BK, WGM, WGN, MINW) \
    case id: emulate<BM, BN, BK, WGM, WGN, V, KA, KB>(p, w); break;
        F4J_FP64_TILES(F4J_TILE_CASE)
#undef F4J_TILE_CASE
        default: CHECK(false, "unknown tile config %d", cfg);
    }
}

template <int KA, int KB>
void run_cfg(int cfg, int vec, const GemmProblem &p, std::vector<int> &w) {
    if (vec == 2) run_v<2, KA, KB>(cfg, p, w);
    else run_v<1, KA, KB>(cfg, p, w);
}

// element (r,c) of op(X) where op(X) is R x Ccols; X complex interleaved or real
static void get(const std::vector<double> &X, int kind, int64_t R, int64_t Cc, int64_t r, int64_t c, double &re,
                double &im) {
    const bool cplx = kind >= Z_N, trans = (kind == R_T || kind == Z_T || kind == Z_C);
    const int64_t idx = trans ? c * R + r : r * Cc + c;
    if (!cplx) {
        re = X[idx];
        im = 0;
        return;
    }
    re = X[2 * idx];
    im = (kind == Z_C) ? -X[2 * idx + 1] : X[2 * idx + 1];
}

int main(int argc, char **argv) {
    if (argc < 7) return 2;
    const int ka = std::atoi(argv[1]), kb = std::atoi(argv[2]);
    const int64_t M = std::atoll(argv[3]), N = std::atoll(argv[4]), K = std::atoll(argv[5]);
    const int cfg = std::atoi(argv[6]);
    const int vec = argc > 7 ? std::atoi(argv[7]) : 1;
    const bool cplx = ka >= Z_N;
    const int w = cplx ? 2 : 1;
    std::mt19937_64 rng(1234 + M * 7 + N * 13 + K * 17 + ka * 3 + kb);
    std::uniform_int_distribution<int> d(-8, 8);
    std::vector<double> A(M * K * w), B(K * N * w);
    for (auto &x : A) x = d(rng);
    for (auto &x : B) x = d(rng);
    const bool ta = (ka == R_T || ka == Z_T || ka == Z_C), tb = (kb == R_T || kb == Z_T || kb == Z_C);
    GemmProblem p;
    p.M = M;
    p.N = N * w;
    p.K = K * w;
    p.A = A.data();
    p.lda = (ta ? M : K) * w;
    p.B = B.data();
    p.ldb = (tb ? K : N) * w;
    std::vector<double> C(M * N * w, -777.0);
    p.C = C.data();
    p.ldc = N * w;
    std::vector<int> written(M * N * w, 0);
    // 16-byte loads only where the library would use them.
    const int vec_used = (vec == 2 && vec2_ok(p.A, p.lda, stored_cols(ka, true, p.M, p.N, p.K)) &&
                          vec2_ok(p.B, p.ldb, stored_cols(kb, false, p.M, p.N, p.K))) ? 2 : 1;
#define CASE(A_, B_) \
    if (ka == A_ && kb == B_) run_cfg<A_, B_>(cfg, vec_used, p, written);
    CASE(R_N, R_N) CASE(R_N, R_T) CASE(R_T, R_N) CASE(R_T, R_T)
    CASE(Z_N, Z_N) CASE(Z_N, Z_T) CASE(Z_N, Z_C) CASE(Z_T, Z_N) CASE(Z_T, Z_T) CASE(Z_T, Z_C)
    CASE(Z_C, Z_N) CASE(Z_C, Z_T) CASE(Z_C, Z_C)
#undef CASE
    for (size_t i = 0; i < written.size(); ++i) CHECK(written[i] == 1, "C[%zu] written %d times", i, written[i]);
    for (int64_t i = 0; i < M && !g_fail; ++i)
        for (int64_t j = 0; j < N; ++j) {
            double sr = 0, si = 0;
            for (int64_t k = 0; k < K; ++k) {
                double ar, ai, br, bi;
                get(A, ka, M, K, i, k, ar, ai);
                get(B, kb, K, N, k, j, br, bi);
                sr += ar * br - ai * bi;
                si += ar * bi + ai * br;
            }
            const double gr = C[(i * N + j) * w], gi = cplx ? C[(i * N + j) * w + 1] : 0.0;
            CHECK(gr == sr && gi == si, "C(%ld,%ld) = (%g,%g) want (%g,%g)", (long)i, (long)j, gr, gi, sr, si);
        }
    if (g_fail) {
        std::printf("FAILED %d\n", g_fail);
        return 1;
    }
    std::printf("OK vec=%d\n", vec_used);
    return 0;
}
