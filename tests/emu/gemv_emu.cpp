// Host-side emulator of the matrix-vector kernels' index math (gemv.hip).
//
// Runs, lane by lane and step by step on the CPU, with the kernels' own index
// functions (flag4j_amd/csrc/gemv_layout.h):
//   dma    gemv_dma_kernel (direct-to-LDS ring): every buffer load writes LDS
//          asynchronously; loads retire in issue order and only when a vmcnt
//          wait lets them (gemv::dma_wait).  Checks: every load in bounds of its
//          buffer resource, no LDS byte read while a load that writes it is in
//          flight, every byte read was written by a load of the stage being
//          summed, and y equals the blocked rule bit for bit.
//   reg    gemv_kernel (register-staged, VEC 1, 3 stages ahead): loads past K or
//          past the slab return zeros; same result check.
//   split  gemv_split_kernel for one kind (chain, cplx, rx, cx): 256 threads form
//          a chunk's products into alternating LDS buffers, chain lanes add the
//          previous chunk's; checks that no product is read in the barrier phase
//          that writes its buffer, that every product is read once, in k order,
//          and that y equals the k-ascending sum bit for bit.
// The references are written here directly from the reference's summation
// rules (RealDenseMatMult.blockedVector; one k-ascending chain per row and
// component, Complex128 products), not from gemv_layout.h.
//
// usage: gemv_emu dma2|dma1|reg <M> <K> <lda> <incx> [seed]
//        gemv_emu split <chain|cplx|rx|cx> <M> <K> [seed]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <random>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../flag4j_amd/csrc/gemv_layout.h"

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

static bool same(double a, double b) { return std::memcmp(&a, &b, 8) == 0; }

// values with many magnitudes, so a wrong order or a wrong element changes the bits
static std::vector<double> values(std::mt19937_64 &rng, size_t n) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::uniform_int_distribution<int> e(-20, 20);
    std::vector<double> v(n);
    for (auto &x : v) x = std::ldexp(u(rng), e(rng));
    return v;
}

// RealDenseMatMult.blockedVector: y_i = sum over 64-wide blocks (in order) of
// the block's k-ascending sum from +0.0, added to a running sum from +0.0.
static std::vector<double> blocked_ref(const std::vector<double> &A, const std::vector<double> &x, int64_t M,
                                       int64_t K, int64_t lda, int64_t incx) {
    std::vector<double> y(M);
    for (int64_t i = 0; i < M; ++i) {
        double dest = 0.0;
        for (int64_t b = 0; b < K; b += 64) {
            double t = 0.0;
            for (int64_t k = b; k < K && k < b + 64; ++k) t = t + A[i * lda + k] * x[k * incx];
            dest = dest + t;
        }
        y[i] = dest;
    }
    return y;
}

// ---- gemv_dma_kernel<GEMV_BLOCKED, XV, D, 8> ----

struct PendingLoad {
    int64_t stage;
    std::vector<std::pair<int, std::vector<unsigned char>>> writes;   // (LDS byte, bytes)
};

template <int XV, int D>
static void run_dma(const std::vector<double> &A, const std::vector<double> &x, int64_t M, int64_t K, int64_t lda,
                    int64_t incx, std::vector<double> &y) {
    constexpr int SLOT = gemv::slot_bytes(), XL = gemv::x_dma_loads<XV>(), L = gemv::dma_loads<XV>();
    const int64_t nstage = gemv::stages_for(K);
    const int64_t waves = (M + gemv::ROWS - 1) / gemv::ROWS;
    for (int64_t w = 0; w < waves; ++w) {
        const int64_t row0 = w * gemv::ROWS;
        const int rows = (int)std::min<int64_t>(gemv::ROWS, M - row0);
        const uint64_t ra_bytes = (uint64_t)rows * lda * 8, rx_bytes = (uint64_t)((K - 1) * incx + 1) * 8;
        const unsigned char *ra = (const unsigned char *)(A.data() + row0 * lda);
        const unsigned char *rx = (const unsigned char *)x.data();
        const int ka_max = (int)K - 2, kx_max = XV == 2 ? (int)K - 2 : (int)K - 1;
        int akl[64][8], xkl[64][4], xadd[64][4];
        uint32_t arow[64][8];
        for (int l = 0; l < 64; ++l) {
            for (int i = 0; i < gemv::DMA_NL; ++i) {
                int r, kl;
                gemv::a_slot<2>(i, l, r, kl);
                akl[l][i] = kl;
                arow[l][i] = (uint32_t)((int64_t)(r < rows ? r : rows - 1) * lda * 8);
            }
            for (int h = 0; h < XL; ++h) gemv::dma_x_slot<XV>(h, l, xkl[l][h], xadd[l][h]);
        }
        std::vector<unsigned char> lds(D * SLOT, 0xCD);
        std::vector<int64_t> tag(D * SLOT, -1);       // stage of the load that last wrote the byte
        std::vector<int> pending_n(D * SLOT, 0);      // loads in flight that write the byte
        std::deque<PendingLoad> q;                    // one entry per wave-wide load instruction
        auto retire_to = [&](size_t keep) {
            while (q.size() > keep) {
                for (auto &wr : q.front().writes) {
                    std::memcpy(&lds[wr.first], wr.second.data(), wr.second.size());
                    for (size_t b = 0; b < wr.second.size(); ++b) {
                        tag[wr.first + b] = q.front().stage;
                        --pending_n[wr.first + b];
                    }
                }
                q.pop_front();
            }
        };
        auto load = [&](int64_t stage, const unsigned char *base, uint64_t nbytes, int lds0, int size,
                        const uint32_t *off) {
            PendingLoad pl{stage, {}};
            for (int l = 0; l < 64; ++l) {
                CHECK(off[l] + (uint64_t)size <= nbytes, "wave %lld stage %lld lane %d: load at byte %u + %d past %llu",
                      (long long)w, (long long)stage, l, off[l], size, (unsigned long long)nbytes);
                std::vector<unsigned char> b(size, 0);
                if (off[l] + (uint64_t)size <= nbytes) std::memcpy(b.data(), base + off[l], size);
                const int dst = lds0 + size * l;
                CHECK(dst >= 0 && dst + size <= D * SLOT, "LDS write out of range");
                for (int k = 0; k < size; ++k) ++pending_n[dst + k];
                pl.writes.push_back({dst, std::move(b)});
            }
            q.push_back(std::move(pl));
        };
        auto issue = [&](int64_t s) {
            const int ks = gemv::kstage(s);
            const int slot = (int)(s % D) * SLOT;
            uint32_t off[64];
            for (int i = 0; i < gemv::DMA_NL; ++i) {
                for (int l = 0; l < 64; ++l) off[l] = arow[l][i] + (uint32_t)std::min(ks + akl[l][i], ka_max) * 8u;
                load(s, ra, ra_bytes, slot + gemv::dma_a_store(i, 0), 16, off);
            }
            for (int h = 0; h < XL; ++h) {
                for (int l = 0; l < 64; ++l)
                    off[l] = (uint32_t)std::min(ks + xkl[l][h], kx_max) * (uint32_t)incx * 8u + xadd[l][h];
                load(s, rx, rx_bytes, slot + gemv::dma_x_store<XV>(h, 0), XV == 2 ? 16 : 4, off);
            }
        };
        auto rd = [&](int64_t s, int byte) {
            for (int b = 0; b < 8; ++b) {
                CHECK(pending_n[byte + b] == 0, "wave %lld stage %lld: LDS byte %d read with a load in flight",
                      (long long)w, (long long)s, byte + b);
                CHECK(tag[byte + b] == s, "wave %lld stage %lld: LDS byte %d holds stage %lld", (long long)w,
                      (long long)s, byte + b, (long long)tag[byte + b]);
            }
            double v;
            std::memcpy(&v, &lds[byte], 8);
            return v;
        };
        double s_acc[64] = {}, dest[64] = {};
        for (int d = 0; d < D - 1; ++d) issue(d);
        for (int64_t s = 0; s < nstage; ++s) {
            retire_to(gemv::dma_wait(D, L));
            issue(s + D - 1);
            const int ks = gemv::kstage(s);
            const bool masked = !(ks + gemv::SPAN <= K);
            const int slot = (int)(s % D) * SLOT;
            for (int l = 0; l < 64; ++l) {
                const int kb = 64 * (l & 7);
                for (int p = 0; p < 8; ++p) {
                    const double a0 = rd(s, slot + gemv::dma_a_read(l, p)), a1 = rd(s, slot + gemv::dma_a_read(l, p) + 8);
                    const double x0 = rd(s, slot + gemv::dma_x_read<XV>(l, p, 0));
                    const double x1 = rd(s, slot + gemv::dma_x_read<XV>(l, p, 1));
                    const double v = (s_acc[l] + a0 * x0) + a1 * x1;
                    if (!masked || ks + kb + 2 * p < K) s_acc[l] = v;
                }
            }
            if ((s & 3) == 3) {
                for (int l = 0; l < 64; ++l) {
                    double t = dest[l];
                    for (int b = 0; b < 8; ++b) t = t + s_acc[(l & ~7) + b];
                    dest[l] = t;
                }
                for (int l = 0; l < 64; ++l) s_acc[l] = 0.0;
            }
        }
        retire_to(0);
        for (int l = 0; l < 64; ++l) {
            const int r = l >> 3;
            if ((l & 7) == 0 && r < rows) y[row0 + r] = dest[l];
        }
    }
}

// ---- gemv_kernel<GEMV_BLOCKED, 1, 3> (register-staged) ----

static void run_reg(const std::vector<double> &A, const std::vector<double> &x, int64_t M, int64_t K, int64_t lda,
                    int64_t incx, std::vector<double> &y) {
    constexpr int D = 3, NA = 16, XN = 2;
    constexpr uint32_t OOR = 0x80000000u;
    const int64_t stages = gemv::stages_for(K), nstage = (stages + D - 1) / D * D;
    const int64_t waves = (M + gemv::ROWS - 1) / gemv::ROWS;
    for (int64_t w = 0; w < waves; ++w) {
        const int64_t row0 = w * gemv::ROWS;
        const int rows = (int)std::min<int64_t>(gemv::ROWS, M - row0);
        const uint64_t ra_bytes = (uint64_t)rows * lda * 8, rx_bytes = (uint64_t)((K - 1) * incx + 1) * 8;
        const double *ra = A.data() + row0 * lda;
        int akl[64][NA], xkl[64][XN];
        uint32_t aoff[64][NA], xoff[64][XN];
        for (int l = 0; l < 64; ++l) {
            for (int i = 0; i < NA; ++i) {
                int r, kl;
                gemv::a_slot<1>(i, l, r, kl);
                akl[l][i] = r < rows ? kl : (int)K;
                aoff[l][i] = (uint32_t)(((int64_t)r * lda + kl) * 8);
            }
            for (int h = 0; h < XN; ++h) {
                xkl[l][h] = gemv::x_slot(h, l);
                xoff[l][h] = (uint32_t)((int64_t)xkl[l][h] * incx * 8);
            }
        }
        auto fetch = [&](const double *base, uint64_t nbytes, uint32_t off, int64_t s, int l) {
            if (off == OOR) return 0.0;
            CHECK(off + 8ull <= nbytes, "wave %lld stage %lld lane %d: load at byte %u past %llu", (long long)w,
                  (long long)s, l, off, (unsigned long long)nbytes);
            return off + 8ull <= nbytes ? base[off / 8] : 0.0;
        };
        double va[D][64][NA], vx[D][64][XN];
        auto issue = [&](int d, int64_t s) {
            const int ks = gemv::kstage(s);
            for (int l = 0; l < 64; ++l) {
                for (int i = 0; i < NA; ++i)
                    va[d][l][i] = fetch(ra, ra_bytes, ks + akl[l][i] < K ? aoff[l][i] + (uint32_t)ks * 8u : OOR, s, l);
                for (int h = 0; h < XN; ++h)
                    vx[d][l][h] = fetch(x.data(), rx_bytes,
                                        ks + xkl[l][h] < K ? xoff[l][h] + (uint32_t)ks * (uint32_t)incx * 8u : OOR, s, l);
            }
        };
        std::vector<double> lds(1024 + 64 * XN);
        std::vector<int> written(lds.size());
        double s_acc[64] = {}, dest[64] = {};
        for (int d = 0; d < D; ++d) issue(d, d);
        for (int64_t s0 = 0; s0 < nstage; s0 += D) {
            for (int d = 0; d < D; ++d) {
                const int64_t s = s0 + d;
                std::fill(written.begin(), written.end(), 0);
                for (int l = 0; l < 64; ++l) {
                    for (int i = 0; i < NA; ++i) {
                        lds[gemv::a_store<1>(i, l)] = va[d][l][i];
                        ++written[gemv::a_store<1>(i, l)];
                    }
                    for (int h = 0; h < XN; ++h) {
                        lds[gemv::x_store(h, l)] = vx[d][l][h];
                        ++written[gemv::x_store(h, l)];
                    }
                }
                issue(d, s + D);
                for (int l = 0; l < 64; ++l)
                    for (int e = 0; e < 16; ++e) {
                        const int ia = gemv::a_read<1>(l, e), ix = gemv::x_read(l, e);
                        CHECK(written[ia] == 1 && written[ix] == 1, "stage %lld lane %d: LDS read of an unwritten or "
                              "twice-written slot", (long long)s, l);
                        s_acc[l] = s_acc[l] + lds[ia] * lds[ix];
                    }
                if ((s & 3) == 3) {
                    for (int l = 0; l < 64; ++l) {
                        double t = dest[l];
                        for (int b = 0; b < 8; ++b) t = t + s_acc[(l & ~7) + b];
                        dest[l] = t;
                    }
                    for (int l = 0; l < 64; ++l) s_acc[l] = 0.0;
                }
            }
        }
        for (int l = 0; l < 64; ++l) {
            const int r = l >> 3;
            if ((l & 7) == 0 && r < rows) y[row0 + r] = dest[l];
        }
    }
}

// ---- gemv_split_kernel<KIND, 16> ----

enum Kind { CHAIN, CPLX, RX, CX };

static void run_split(Kind kind, const std::vector<double> &A, const std::vector<double> &x, int64_t M, int64_t K,
                      std::vector<double> &y) {
    constexpr int R = 16, KC = 64, E = KC * R / 256, TPR = KC / E;
    const bool AC = kind == CPLX || kind == CX, XC = kind == CPLX || kind == RX;
    const int C = (AC || XC) ? 2 : 1, aw = AC ? 2 : 1, xw = XC ? 2 : 1;
    const int64_t lda = K * aw;                      // doubles per row (unpadded here)
    const int64_t nchunk = (K + KC - 1) / KC;
    const int64_t blocks = (M + R - 1) / R;
    for (int64_t blk = 0; blk < blocks; ++blk) {
        const int64_t row0 = blk * R;
        // prod[buf][row][k][comp], the barrier phase that wrote it, and reads
        std::vector<double> prod(2 * R * (KC + 1) * 2, 0.0);
        std::vector<int> wphase(prod.size(), -1), reads(prod.size(), 0);
        int phase = 0;
        double av[256][E][2], xv[256][E][2];
        auto load = [&](int64_t c) {
            for (int t = 0; t < 256; ++t) {
                const int r = t / TPR, q = t % TPR;
                const int64_t row = row0 + r;
                for (int u = 0; u < E; ++u) {
                    const int64_t k = c * KC + E * q + u;
                    const bool ok = row < M && k < K;
                    for (int cc = 0; cc < 2; ++cc) {
                        av[t][u][cc] = ok && cc < aw ? A[row * lda + aw * k + cc] : 0.0;
                        xv[t][u][cc] = k < K && cc < xw ? x[xw * k + cc] : 0.0;
                    }
                }
            }
        };
        auto idx = [&](int buf, int r, int k, int comp) { return ((buf * R + r) * (KC + 1) + k) * 2 + comp; };
        auto products = [&](int buf) {
            for (int t = 0; t < 256; ++t) {
                const int r = t / TPR, q = t % TPR;
                for (int u = 0; u < E; ++u) {
                    const double ar = av[t][u][0], ai = av[t][u][1], xr = xv[t][u][0], xi = xv[t][u][1];
                    double v0, v1 = 0.0;
                    if (kind == CPLX) {
                        v0 = ar * xr - ai * xi;
                        v1 = ar * xi + ai * xr;
                    } else if (kind == RX) {
                        v0 = xr * ar;
                        v1 = xi * ar;
                    } else if (kind == CX) {
                        v0 = ar * xr;
                        v1 = ai * xr;
                    } else {
                        v0 = ar * xr;
                    }
                    const int k = E * q + u;
                    prod[idx(buf, r, k, 0)] = v0;
                    prod[idx(buf, r, k, 1)] = v1;
                    wphase[idx(buf, r, k, 0)] = wphase[idx(buf, r, k, 1)] = phase;
                }
            }
        };
        double sacc[R][2] = {};
        load(0);
        products(0);
        if (nchunk > 1) load(1);
        ++phase;                                       // __syncthreads
        for (int64_t c = 0; c < nchunk; ++c) {
            if (c + 1 < nchunk) {
                products((int)((c + 1) & 1));
                if (c + 2 < nchunk) load(c + 2);
            }
            for (int t = 0; t < C * R; ++t) {
                const int cr = t % R, comp = t / R;
                const int kn = (int)std::min<int64_t>(KC, K - c * KC);
                for (int k = 0; k < kn; ++k) {
                    const int i = idx((int)(c & 1), cr, k, comp);
                    CHECK(wphase[i] >= 0 && wphase[i] < phase, "block %lld chunk %lld: product read in the phase "
                          "that writes it (or never written)", (long long)blk, (long long)c);
                    ++reads[i];
                    sacc[cr][comp] = sacc[cr][comp] + prod[i];
                }
            }
            ++phase;                                   // __syncthreads
        }
        for (int t = 0; t < C * R; ++t) {
            const int cr = t % R, comp = t / R;
            if (row0 + cr < M) y[(row0 + cr) * C + comp] = sacc[cr][comp];
        }
    }
}

static std::vector<double> split_ref(Kind kind, const std::vector<double> &A, const std::vector<double> &x, int64_t M,
                                     int64_t K) {
    const bool AC = kind == CPLX || kind == CX, XC = kind == CPLX || kind == RX;
    const int C = (AC || XC) ? 2 : 1;
    std::vector<double> y(M * C);
    for (int64_t i = 0; i < M; ++i) {
        double re = 0.0, im = 0.0;
        for (int64_t k = 0; k < K; ++k) {
            if (kind == CPLX) {            // Complex128.mult: (ar xr - ai xi, ar xi + ai xr)
                const double ar = A[i * 2 * K + 2 * k], ai = A[i * 2 * K + 2 * k + 1], xr = x[2 * k], xi = x[2 * k + 1];
                re = re + (ar * xr - ai * xi);
                im = im + (ar * xi + ai * xr);
            } else if (kind == RX) {       // real a times complex x
                re = re + x[2 * k] * A[i * K + k];
                im = im + x[2 * k + 1] * A[i * K + k];
            } else if (kind == CX) {       // complex a times real x
                re = re + A[i * 2 * K + 2 * k] * x[k];
                im = im + A[i * 2 * K + 2 * k + 1] * x[k];
            } else {
                re = re + A[i * K + k] * x[k];
            }
        }
        y[i * C] = re;
        if (C == 2) y[i * C + 1] = im;
    }
    return y;
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "split") {
        if (argc < 5) return 2;
        const std::string kn = argv[2];
        const Kind kind = kn == "cplx" ? CPLX : kn == "rx" ? RX : kn == "cx" ? CX : CHAIN;
        const int64_t M = std::atoll(argv[3]), K = std::atoll(argv[4]);
        std::mt19937_64 rng(argc > 5 ? std::atoll(argv[5]) : 1);
        const int aw = (kind == CPLX || kind == CX) ? 2 : 1, xw = (kind == CPLX || kind == RX) ? 2 : 1;
        const auto A = values(rng, M * K * aw), x = values(rng, K * xw);
        const int C = (aw == 2 || xw == 2) ? 2 : 1;
        std::vector<double> y(M * C, NAN);
        run_split(kind, A, x, M, K, y);
        const auto ref = split_ref(kind, A, x, M, K);
        for (int64_t i = 0; i < M * C; ++i) CHECK(same(y[i], ref[i]), "y[%lld] differs from the chain", (long long)i);
    } else {
        if (argc < 6) return 2;
        const int64_t M = std::atoll(argv[2]), K = std::atoll(argv[3]), lda = std::atoll(argv[4]),
                      incx = std::atoll(argv[5]);
        std::mt19937_64 rng(argc > 6 ? std::atoll(argv[6]) : 1);
        const auto A = values(rng, M * lda), x = values(rng, (K - 1) * incx + 1);
        std::vector<double> y(M, NAN);
        if (mode == "dma2") {
            CHECK(K % 2 == 0 && lda % 2 == 0 && incx == 1, "dma2 needs even K and lda and unit incx");
            run_dma<2, 4>(A, x, M, K, lda, incx, y);
        } else if (mode == "dma1") {
            CHECK(K % 2 == 0 && lda % 2 == 0, "dma1 needs even K and lda");
            run_dma<1, 4>(A, x, M, K, lda, incx, y);
        } else {
            run_reg(A, x, M, K, lda, incx, y);
        }
        const auto ref = blocked_ref(A, x, M, K, lda, incx);
        for (int64_t i = 0; i < M; ++i) CHECK(same(y[i], ref[i]), "y[%lld] differs from the blocked rule", (long long)i);
    }
    if (g_fail) {
        std::printf("%d failures\n", g_fail);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
