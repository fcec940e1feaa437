// Batched split-precision GEMM: many independent real or complex products,
// each C_p = op(A_p) op(B_p), in a fixed sequence of launches (exponents,
// digit slices, one persistent slice-product kernel over every output tile of
// every product, residual check) that can be captured into a hipGraph.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace f4j {

struct BatchProblem {
    int64_t M, N, K;
    const double *A;      // leading dimensions in elements (complex: interleaved re, im pairs)
    int64_t lda;
    const double *B;
    int64_t ldb;
    double *C;
    int64_t ldc;
};

struct SplitBatch;

// Plan a batch on `carrier` (SPLIT_BF16 / SPLIT_FP16) with `slices` digits per
// operand, or on SPLIT_INT8 (the modular form) with `slices` moduli (0: the
// fewest that give t >= 64 bits at the batch's largest K).  Complex products
// use the 3M form (three real products) when use3m, else four.  `check_rows`
// rows of every product are verified by the residual check (0 = off).  On
// failure returns nullptr with *status and *err set.
SplitBatch *split_batch_create(bool cplx, int opA, int opB, int carrier, int slices, bool use3m, int check_rows,
                               const std::vector<BatchProblem> &probs, std::string *err, int *status);
// Enqueue the whole batch on `stream` (no allocation, no memset: capturable).
// ev_begin / ev_end (optional) are recorded around the slice-product launch.
hipError_t split_batch_enqueue(const SplitBatch *b, hipStream_t stream, hipEvent_t ev_begin = nullptr,
                               hipEvent_t ev_end = nullptr);
// Per-problem max |C - AB| / (|A||B|) over the checked rows of the last run
// (-1 where nothing was checked; NaN where the problem's operands hold NaN or
// Inf, which the digit slices cannot carry); synchronises `stream`.
hipError_t split_batch_errors(const SplitBatch *b, double *out, int64_t count, hipStream_t stream);
// Device array of the last run's per-product error bits (the bits of the
// non-negative error, or a NaN pattern above every finite error for NaN / Inf
// operands), and the slot of each problem in it (-1: no product computed).
const unsigned long long *split_batch_err_bits(const SplitBatch *b);
const std::vector<int> &split_batch_slots(const SplitBatch *b);
int split_batch_real_products(const SplitBatch *b);     // real slice products per run (jobs x pairs, or jobs x moduli)
int split_batch_checked_rows(const SplitBatch *b);
size_t split_batch_bytes(const SplitBatch *b);          // device memory held
void split_batch_destroy(SplitBatch *b);

}  // namespace f4j
