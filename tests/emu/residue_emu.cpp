// Host check of the modular GEMM's residue arithmetic (flag4j_amd/csrc/crt_residue.h)
// against exact integer residues.
//
// Operand residues: for scaled integers v with |v| <= 2^t, t = 1 .. 75, the
// residue returned for each of the 20 moduli must be congruent to v and lie in
// [-128, 127] (m = 256 may also give 128, the same byte).  Inputs: random
// magnitudes at every t, the extremes +-2^t, and the cases where the quotient
// (u + 128) / m is an exact integer (v = k m - 128), the only ones in which the
// kernel's floor may land one below.
// The biased form of the split kernels (residue_biased, pack4_biased) must give
// the word r + 128 and pack to the same bytes, and their per-line shift
// (scaled_shift, scaled_int_sh) the same scaled integers as scaled_int.
// Product residues: random int32 accumulators, multiples of m and the int32
// extremes; the result must be congruent and lie in [0, m].
// Prints "OK <count>" or the first mismatches and exits non-zero.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <random>

#define F4J_HD inline
#include "../../flag4j_amd/csrc/crt_residue.h"

static const int MODULI[20] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                               223, 217, 211, 199, 197, 193, 191, 181, 179, 173};

// exact v mod m in [0, m) for an integer-valued double
static int exact_mod(double v, int m) {
    if (v == 0.0) return 0;
    int e;
    const double f = std::frexp(v, &e);            // v = f 2^e, |f| in [0.5, 1)
    __int128 mant = (__int128)std::ldexp(f, 53);   // v = mant 2^(e - 53)
    const int sh = e - 53;
    __int128 x;
    if (sh >= 0) {
        long long p = 1 % m;
        for (int i = 0; i < sh; ++i) p = (p * 2) % m;
        x = (mant % m) * p;
    } else {
        x = mant / ((__int128)1 << (-sh));            // v is an integer: exact
    }
    int r = (int)(x % m);
    return r < 0 ? r + m : r;
}

static long bad = 0, total = 0;

static void check_operand(double v, int t) {
    const bool wide = t > 58;
    double vh = 0, vl = 0;
    if (!wide) f4j::crt::residue_split(v, vh, vl);
    for (int l = 0; l < 20; ++l) {
        const int m = MODULI[l];
        const double mi = 1.0 / m, c26 = (double)((1 << 26) % m);
        const int r = wide ? f4j::crt::residue_wide(v, m, mi, c26) : f4j::crt::residue_of(vh, vl, m, mi, c26);
        const int want = exact_mod(v, m);
        const bool ok = ((r % m) + m) % m == want && r >= -128 && (r <= 127 || (m == 256 && r == 128));
        ++total;
        if (!ok && bad++ < 10) std::printf("operand t=%d v=%.17g m=%d got %d want %d (mod m)\n", t, v, m, r, want);
        // the split kernels' biased form: word r + 128, and the packed byte equals the byte of r
        const double cq = f4j::crt::residue_cq(mi);
        const unsigned rb = wide ? f4j::crt::residue_wide_biased(v, m, mi, cq, c26)
                                 : f4j::crt::residue_biased(vh, f4j::crt::residue_bias(vl), m, mi, cq, c26);
        const unsigned word = f4j::crt::pack4_biased(rb, 0u, rb, 128u);
        const bool okb = rb == (unsigned)(r + 128) && (word & 255u) == ((unsigned)r & 255u) &&
                         ((word >> 16) & 255u) == ((unsigned)r & 255u) && ((word >> 8) & 255u) == 128u &&
                         (word >> 24) == 0u;
        ++total;
        if (!okb && bad++ < 10)
            std::printf("biased t=%d v=%.17g m=%d got word %u (r %d) packed %08x\n", t, v, m, rb, r, word);
    }
}

int main() {
    std::mt19937_64 g(7);
    for (int t = 1; t <= 75; ++t) {
        const double top = std::ldexp(1.0, t);
        check_operand(top, t);
        check_operand(-top, t);
        for (int it = 0; it < 4000; ++it) {
            const double u = std::ldexp((double)(g() >> 11), -53) * 2 - 1;                 // [-1, 1)
            const double sc = (it % 4 == 0) ? std::ldexp(1.0, -(int)(g() % 40)) : 1.0;
            check_operand(std::rint(std::ldexp(u * sc, t)), t);
        }
        // (u + 128) / m an exact integer: v = k m - 128 (u = v mod m matches)
        for (int l = 0; l < 20 && t > 9; ++l) {
            const int m = MODULI[l];
            for (int it = 0; it < 200; ++it) {
                const double k = std::floor(std::ldexp((double)(g() >> 11), -53) * std::ldexp(1.0, t - 9));
                const double v = std::fma(k, (double)m, -128.0);
                if (std::fabs(v) <= top) {
                    check_operand(v, t);
                    check_operand(-v - 256.0, t);
                }
            }
        }
    }
    // per-line shift of the split kernels: the same scaled integers, zeros for an all-zero line
    for (int it = 0; it < 200000; ++it) {
        const int t = 1 + (int)(g() % 75);
        const double x = std::ldexp((double)(g() >> 11), -53) * 2 - 1;
        int e;
        (void)std::frexp(x * std::ldexp(1.0, (int)(g() % 2000) - 1000), &e);
        const double v = std::ldexp(x, e - 1);                 // |v| < 2^e
        const int em = e + (int)(g() % 4);                     // the line's exponent, >= e
        const double a = f4j::crt::scaled_int(v, em, t), b = f4j::crt::scaled_int_sh(v, f4j::crt::scaled_shift(em, t));
        const double z = f4j::crt::scaled_int_sh(0.0, f4j::crt::scaled_shift(INT_MIN, t));
        ++total;
        if (!(a == b || (std::isnan(a) && std::isnan(b))) || z != 0.0 || std::signbit(z)) {
            if (bad++ < 10) std::printf("shift t=%d v=%.17g e=%d: %.17g vs %.17g, zero line %.17g\n", t, v, em, a, b, z);
        }
    }
    // product residues
    for (int l = 0; l < 20; ++l) {
        const int m = MODULI[l];
        const double mi = 1.0 / m;
        auto chk = [&](long long a) {
            if (a < INT32_MIN || a > INT32_MAX) return;
            const int r = f4j::crt::product_residue((int)a, m, mi);
            const long long want = ((a % m) + m) % m;
            ++total;
            if (!(r >= 0 && r <= m && r % m == want) && bad++ < 10)
                std::printf("product a=%lld m=%d got %d want %lld\n", a, m, r, want);
        };
        chk(INT32_MAX);
        chk(INT32_MIN);
        chk(0);
        for (int it = 0; it < 20000; ++it) {
            const long long a = (long long)(int32_t)(uint32_t)g();
            chk(a);
            chk(a / m * m);                  // exact multiples
        }
    }
    if (bad) {
        std::printf("FAIL %ld of %ld\n", bad, total);
        return 1;
    }
    std::printf("OK %ld\n", total);
    return 0;
}
