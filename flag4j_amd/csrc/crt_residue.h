// Residue arithmetic of the modular (CRT) GEMM, shared by its kernels and by
// the host test that checks it against exact integer residues
// (tests/emu/residue_emu.cpp).
//
// Operand residues.  v is an integer-valued double (a scaled operand entry,
// |v| <= 2^t).  For |v| < 2^58 it is split once, exactly, into
//   v = vh 2^26 + vl,  vh = floor(v 2^-26),  vl in [0, 2^26),
// and for each modulus m (c26 = 2^26 mod m, minv = 1/m rounded)
//   u = vh c26 + vl           exact: |u| < 2^41, u = v (mod m)
//   q = floor((u + 128) / m)  from one fma: the computed quotient is within
//                             2^-19 of (u + 128) / m, whose fractional part is a
//                             multiple of 1/m >= 2^-8, so q is exact unless
//                             (u + 128) / m is an integer, where q may be one less;
//   r = u - q m               exact, r = v (mod m), r in [-128, m - 129], or
//                             r = m - 128 in that integer case.
// So r always lies in [-128, 127] (m <= 255), or [-128, 128] for m = 256, whose
// 128 is the byte of -128: every residue fits a signed byte with no correction.
// Wider v (2^58 <= |v| <= 2^75) is first reduced to |v| < 2^24 by one exact step
// v - rint(v / m) m (the computed quotient is off by at most 2^16).
// The split kernels evaluate the same steps biased by 2^52 + 128 (residue_biased):
//   U = vh c26 + (vl + 2^52 + 128)   exact, an integer in (2^52 - 2^41, 2^52 + 2^41);
//   q = floor(U minv - 2^52 minv)    one fma: 2^52 minv is exact, so this is the
//                                     same rounding of (u + 128) / m as above, the same q;
//   R = U - q m = 2^52 + r + 128     exact, in [2^52, 2^52 + 256], so the low word of
//                                     R's bits is r + 128 with no conversion;
// and four such words pack into the four signed bytes r with one XOR of 0x80 per
// byte ((r + 128) mod 256 XOR 128 is the byte of r for r in [-128, 128]).
//
// Product residues.  For an int32 accumulator a, r = a - floor(a / m) m with the
// quotient from one multiply (error below 2^-28, fractional parts multiples of
// 1/m): r in [0, m), or r = m when a is a multiple of m.  The value m is kept:
// the reconstruction only uses r modulo m (m W_l = 0 modulo M).
#pragma once
#include <climits>
#include <cmath>
#include <cstdint>

#ifndef F4J_HD
#define F4J_HD __host__ __device__ inline
#endif

namespace f4j {
namespace crt {

constexpr double TWO26 = 67108864.0;

F4J_HD void residue_split(double v, double &vh, double &vl) {
    vh = floor(ldexp(v, -26));
    vl = fma(-vh, TWO26, v);
}

F4J_HD int residue_of(double vh, double vl, double m, double minv, double c26) {
    const double u = fma(vh, c26, vl);
    const double q = floor(fma(u, minv, 128.0 * minv));
    return (int)fma(-q, m, u);
}

// 2^58 <= |v| <= 2^75: one exact reduction, then the small-value step.
F4J_HD int residue_wide(double v, double m, double minv, double c26) {
    const double v1 = fma(-rint(v * minv), m, v);
    return residue_of(0.0, v1, m, minv, c26);
}

// Biased form (header comment): vb = residue_bias(vl) once per entry, cq =
// residue_cq(minv) once per modulus; the result is r + 128, r = residue_of(vh, vl, ...).
constexpr double BIAS = 4503599627370496.0 + 128.0;         // 2^52 + 128
F4J_HD double residue_bias(double vl) { return vl + BIAS; }
F4J_HD double residue_cq(double minv) { return -4503599627370496.0 * minv; }
F4J_HD unsigned residue_biased(double vh, double vb, double m, double minv, double cq, double c26) {
    const double U = fma(vh, c26, vb);
    const double q = floor(fma(U, minv, cq));
    return (unsigned)__builtin_bit_cast(uint64_t, fma(-q, m, U));
}
F4J_HD unsigned residue_wide_biased(double v, double m, double minv, double cq, double c26) {
    const double v1 = fma(-rint(v * minv), m, v);
    return residue_biased(0.0, residue_bias(v1), m, minv, cq, c26);
}
// Four biased residues (r + 128, first in the lowest byte) as the signed bytes r:
// three byte permutes and one XOR on the device.
F4J_HD unsigned pack4_biased(unsigned a, unsigned b, unsigned c, unsigned d) {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned ab = __builtin_amdgcn_perm(b, a, 0x0c0c0400u);     // bytes a0 b0 0 0
    const unsigned cd = __builtin_amdgcn_perm(d, c, 0x0c0c0400u);     // bytes c0 d0 0 0
    return __builtin_amdgcn_perm(cd, ab, 0x05040100u) ^ 0x80808080u;
#else
    return ((a & 255u) | ((b & 255u) << 8) | ((c & 255u) << 16) | (d << 24)) ^ 0x80808080u;
#endif
}

// Product residue in [0, m], as a double (exact integer) and as an int.
F4J_HD double product_residue_d(int a, double m, double minv) {
    const double d = (double)a;
    return fma(-floor(d * minv), m, d);
}
F4J_HD int product_residue(int a, double m, double minv) { return (int)product_residue_d(a, m, minv); }

// Entry v of a line with exponent e (|v| < 2^e; INT_MIN for an all-zero line)
// as a t-bit integer: rint(v 2^(t - e)), |result| <= 2^t.
F4J_HD double scaled_int(double v, int e, int t) { return e == INT_MIN ? 0.0 : rint(ldexp(v, t - e)); }
// The same with the shift taken once per line: scaled_shift(e, t) is t - e, with e
// raised to -4096 for an all-zero line (INT_MIN; its entries are zeros, which any
// shift keeps, and every finite nonzero double has e >= -1073).  A max rather
// than a test for INT_MIN, so that no select is applied per entry.
F4J_HD int scaled_shift(int e, int t) { return t - (e > -4096 ? e : -4096); }
F4J_HD double scaled_int_sh(double v, int sh) { return rint(ldexp(v, sh)); }

// Four residues as the bytes of one 32-bit word (first in the lowest byte).
F4J_HD unsigned pack4(int a, int b, int c, int d) {
    return (unsigned)(a & 255) | ((unsigned)(b & 255) << 8) | ((unsigned)(c & 255) << 16) | ((unsigned)d << 24);
}

}  // namespace crt
}  // namespace f4j
