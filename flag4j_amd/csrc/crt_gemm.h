// Modular (CRT) split-precision GEMM on the INT8 matrix cores: plan and launch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace f4j {

constexpr int CRT_MAX_MODULI = 20;
constexpr int CRT_MAX_CHUNKS = 4;
constexpr int CRT_MAX_JOBS = 4;          // real products of one call: 1 (real), 3 (complex 3M) or 4 (4M)

// Reconstruction constants.  With M = prod m_l and W_l = (M/m_l) * ((M/m_l)^-1 mod m_l),
// X = sum_l r_l W_l = A'B' (mod M).  Every W_l (and M) is cut into P chunks of 40 bits on a
// common grid: W_l = sum_p w[l][p] scale[p], so S_p = sum_l r_l w[l][p] is exact in FP64.
struct CrtTables {
    int n;                               // moduli used
    int t;                               // bits of the scaled integer operands (|A'|, |B'| <= 2^t)
    int P;                               // 40-bit chunks
    int m[CRT_MAX_MODULI];
    double minv[CRT_MAX_MODULI];         // 1 / m_l rounded
    double c26[CRT_MAX_MODULI];          // 2^26 mod m_l
    double w[CRT_MAX_MODULI][CRT_MAX_CHUNKS];
    double mu[CRT_MAX_CHUNKS];           // chunks of M
    double scale[CRT_MAX_CHUNKS];        // 2^(exponent of chunk p)
    double c[CRT_MAX_CHUNKS];            // scale[p] / M
};

struct CrtPlan {
    CrtTables tb{};
    int64_t Mp = 0, Np = 0, Kp = 0;      // padded sizes of the residue slices
    int64_t plane = 0;                   // bytes of one modulus's product residues (Mp Np, crt::res_block layout)
    size_t off_ra = 0, off_rb = 0, off_res = 0, off_ea = 0, off_eb = 0, off_err = 0, off_part = 0, bytes = 0;
    size_t off_flag = 0;                 // int: set by the exponent kernels when an operand holds NaN or Inf
};

// C' of one entry from its chunk sums S_p = sum_l r_l w[l][p] (exact FP64
// integers): q = rint(sum_p S_p c_p) is the multiple of M to remove and
// D_p = S_p - q mu_p are exact; D_p scale_p are exact too, and the lower chunks
// are summed in double-double (TwoSum) because they cancel against the top
// chunk, so C' is rounded once, relative to itself.
static_assert(CRT_MAX_CHUNKS == 4, "crt_recombine takes four chunk sums");
__device__ inline double crt_recombine(double s0, double s1, double s2, double s3, const CrtTables &tb) {
    const double S[CRT_MAX_CHUNKS] = {s0, s1, s2, s3};
    double approx = 0.0;
#pragma unroll
    for (int p = 0; p < CRT_MAX_CHUNKS; ++p) approx = fma(S[p], tb.c[p], approx);
    const double q = rint(approx);
    double hi = 0.0, lo = 0.0;
#pragma unroll
    for (int p = CRT_MAX_CHUNKS - 1; p >= 1; --p) {
        const double v = fma(-q, tb.mu[p], S[p]) * tb.scale[p];
        const double sm = hi + v, z = sm - hi;
        lo += (hi - (sm - z)) + (v - z);
        hi = sm;
    }
    const double top = fma(-q, tb.mu[0], S[0]) * tb.scale[0];
    const double sm = top + hi, z = sm - top;
    return sm + (((top - (sm - z)) + (hi - z)) + lo);
}

// Bits t of the scaled integer operands for n moduli and inner size K: the
// largest t with K 2^(2t) < M/4 (below 1: too few moduli for this K).
int crt_bits(int n, int64_t K);

// Reconstruction tables for n moduli with t = crt_bits(n, K); false if n is
// outside [2, CRT_MAX_MODULI] or t < 1.
bool crt_tables(int n, int64_t K, CrtTables *tb);

// Smallest number of moduli whose product gives t >= 53 bits for inner size K
// (every entry that is within a factor 2 of its line's largest keeps all its bits).
int crt_default_moduli(int64_t K);
// Fewest moduli whose t = crt_bits(n, K) reaches `bits` (CRT_MAX_MODULI if none).
int crt_moduli_for_bits(int bits, int64_t K);

// Plan for C (M x N) = op(A) op(B) with n moduli; returns false if n is
// outside [2, CRT_MAX_MODULI] or gives fewer than 1 bit for this K.
bool plan_crt(int64_t M, int64_t N, int64_t K, int n, int check_rows, CrtPlan *plan);

// A real operand of the products as the kernels read it: line x (row of op(A),
// column of op(B)), element k is ca re + cb im of X at x sx + k sk doubles
// (real data: the value itself).  `lines`: consecutive k are adjacent.
struct CrtOperand {
    const double *X;
    int64_t sx, sk;
    double ca, cb;
    int cplx, lines;
};
// Operand view of X (leading dimension ld in elements; complex X interleaved).
CrtOperand crt_operand(const double *X, int64_t ld, bool lines, bool cplx, double ca = 1.0, double cb = 0.0);

// Where the recombined product goes: real C (C = AB), or the parts of an
// interleaved complex C (Re C (+)= cre AB, Im C (+)= cim AB; first: assign).
struct CrtOutput {
    bool cplx = false;
    double cre = 1.0, cim = 0.0;
    bool first = true;
};

// A second stream for B's exponents and residues: `fork` is recorded on the
// main stream and waited for by `stream`, `join` recorded after B's split and
// waited for by the main stream before the products.
struct CrtSide {
    hipStream_t stream;
    hipEvent_t fork, join;
};

// The real products of a complex call, all on the storage of the same A and B:
// job j multiplies the views a_ca[j] re + a_cb[j] im of op(A) and b_ca[j] re +
// b_cb[j] im of op(B) and contributes cre[j] and cim[j] times its product to
// Re C and Im C.
struct CrtJobs {
    int n = 0;
    double a_ca[CRT_MAX_JOBS] = {}, a_cb[CRT_MAX_JOBS] = {}, b_ca[CRT_MAX_JOBS] = {}, b_cb[CRT_MAX_JOBS] = {};
    double cre[CRT_MAX_JOBS] = {}, cim[CRT_MAX_JOBS] = {};
};
// Bytes between consecutive jobs' workspaces (plan.bytes rounded up to 4 KiB).
size_t crt_job_stride(const CrtPlan &pl);

// Enqueue exponents, residue slices, the n residue products and the FP64
// reconstruction into C.  `ws` holds plan.bytes bytes.
// ev_begin / ev_end (optional) are recorded on `stream` around the product launch.
// With `side`, B's exponents and residues run on side->stream beside A's.
hipError_t launch_crt_gemm(const CrtPlan &pl, const CrtOperand &oa, const CrtOperand &ob, int64_t M, int64_t N,
                           int64_t K, double *C, int64_t ldc, const CrtOutput &out, void *ws, hipStream_t stream,
                           hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr,
                           const CrtSide *side = nullptr);
// All jobs (3 or 4) of a complex product in one sequence: each operand pass (exponents,
// residues) reads the operand once for every job's view, one product launch covers
// every job's moduli and one reconstruction writes the interleaved complex C once.
// oa and ob give the storage (their ca / cb are ignored); `ws` holds
// jobs.n * crt_job_stride(plan) bytes, job j's workspace at j * crt_job_stride(plan).
hipError_t launch_crt_gemm_jobs(const CrtPlan &pl, const CrtOperand &oa, const CrtOperand &ob, int64_t M, int64_t N,
                                int64_t K, double *C, int64_t ldc, const CrtJobs &jobs, void *ws, hipStream_t stream,
                                hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr,
                                const CrtSide *side = nullptr);

}  // namespace f4j
