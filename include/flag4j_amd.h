/*
 * flag4j_amd: MI355X-native dense matrix multiply behind a C ABI.
 *
 * Drop-in backend for Flag4j's dense matmul path.  Citation prefix
 * M: = src/main/java/org/flag4j/ in the Flag4j sources.  Each entry point
 * names the reference interface it replaces.
 *
 * Conventions
 *  - All matrices are row-major.  Leading dimensions (lda, ldb, ldc) are in
 *    elements (complex elements for the z* calls) and must be >= the stored
 *    row length.  Complex data is interleaved (re, im) binary64 pairs.
 *  - op(A) is M x K and op(B) is K x N.  opA == F4J_OP_N means A is stored
 *    M x K; F4J_OP_T / F4J_OP_C mean A is stored K x M (and, for C, its
 *    entries are conjugated).  Same for B (stored K x N, or N x K).
 *  - The library never throws or aborts: every call returns an f4j_status;
 *    f4j_last_error() gives the calling thread's last message.  On failure no
 *    partial result is promised.
 *  - Host-array calls are synchronous and thread-safe; they stage through
 *    library-owned device buffers that are grown geometrically and reused.
 *  - *_dev calls take device pointers and run asynchronously on `stream`
 *    (a hipStream_t; NULL selects the library's own non-blocking stream).
 *  - The first call initialises device 0 if f4j_init was not called.
 */
#ifndef FLAG4J_AMD_H
#define FLAG4J_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define F4J_ABI_VERSION 4

typedef enum f4j_status {
    F4J_OK = 0,
    F4J_ERR_INVALID_ARG = -1,     /* bad op/precision code, null pointer, negative size, ld too small */
    F4J_ERR_SHAPE = -2,           /* inner dimensions differ (Flag4j: LinearAlgebraException) */
    F4J_ERR_OUT_OF_MEMORY = -3,
    F4J_ERR_HIP = -4,             /* a HIP runtime call failed */
    F4J_ERR_NO_DEVICE = -5,
    F4J_ERR_UNSUPPORTED = -6,     /* valid request this build does not implement */
    F4J_ERR_ACCURACY = -7,        /* split-precision residual check failed */
    F4J_ERR_SINGULAR = -8         /* least squares: R is singular (Flag4j: SingularMatrixException) */
} f4j_status;

typedef enum f4j_op { F4J_OP_N = 0, F4J_OP_T = 1, F4J_OP_C = 2 } f4j_op;

typedef enum f4j_precision {
    F4J_PREC_AUTO = -1,           /* dispatcher decides (small -> STRICT, else the configured default) */
    F4J_PREC_FP64 = 0,            /* FP64 matrix cores, FMA accumulation */
    F4J_PREC_STRICT = 1,          /* Flag4j's arithmetic: per element, k ascending, separate mul and add:
                                     bitwise equal to the reference CPU kernels */
    F4J_PREC_SPLIT_BF16 = 2,      /* FP64 split into BF16 slices, FP32-accumulated MFMA, FP64 recombination */
    F4J_PREC_SPLIT_FP16 = 3,      /* same with FP16 slices */
    F4J_PREC_SPLIT_INT8 = 4       /* FP64 as integers modulo coprime moduli <= 256 (Chinese remainder theorem):
                                     INT8 residue products with exact INT32 accumulation, FP64 reconstruction;
                                     `slices` = number of moduli (2..20, 0 = enough for FP64 accuracy) */
} f4j_precision;

typedef struct f4j_device_info {
    int device;
    int compute_units;
    int clock_khz;
    int lds_bytes_per_cu;
    int xcds;
    int64_t hbm_bytes;
    double peak_fp64_tflops;      /* dense FP64 matrix peak used for roofline fractions */
    double peak_bf16_tflops;
    double peak_hbm_gbs;
    char name[64];
    char arch[32];
} f4j_device_info;

/* ---- lifecycle (no reference counterpart: Flag4j has no native layer; the
 *      closest is its lazy static initialisation, M:concurrency/ThreadManager.java:66-88) ---- */
int         f4j_abi_version(void);
int         f4j_init(int device);
int         f4j_shutdown(void);
const char *f4j_last_error(void);
int         f4j_device_info_get(f4j_device_info *out);

/* ---- device memory (device-resident operands, pinned and reused by callers) ---- */
int f4j_dev_alloc(size_t bytes, void **dptr);
int f4j_dev_free(void *dptr);
int f4j_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int f4j_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int f4j_stream_sync(void *stream);

/* ---- real FP64 GEMM ----
 * Replaces, by (opA, opB):
 *   (N, N): Matrix.mult(Matrix)           M:arrays/dense/Matrix.java:638-640 ->
 *           ReDeMatMultDispatcher.dispatch M:linalg/ops/dispatch/ReDeMatMultDispatcher.java:157-237 ->
 *           RealDenseMatMult.*             M:linalg/ops/dense/real/RealDenseMatMult.java:50-415
 *   (N, T): Matrix.multTranspose(Matrix)  M:arrays/dense/Matrix.java:655-663 ->
 *           RealDenseMatrixMultiplyDispatcher.dispatchTranspose M:linalg/ops/RealDenseMatrixMultiplyDispatcher.java:148-154 ->
 *           RealDenseMatMultTranspose.*   M:linalg/ops/dense/real/RealDenseMatMultTranspose.java:53-233
 *   (T, *): A.T().mult(B) without materialising A.T() (M:linalg/ops/TransposeDispatcher.java:77-99).
 * C (M x N) = op(A) op(B).  `slices` is used by the SPLIT precisions (0 = choose): digit slices
 * per operand for SPLIT_BF16/FP16 (1..8), moduli for SPLIT_INT8 (2..20). */
int f4j_dgemm(int opA, int opB, int64_t M, int64_t N, int64_t K,
              const double *A, int64_t lda, const double *B, int64_t ldb,
              double *C, int64_t ldc, int precision, int slices);
int f4j_dgemm_dev(int opA, int opB, int64_t M, int64_t N, int64_t K,
                  const double *A, int64_t lda, const double *B, int64_t ldb,
                  double *C, int64_t ldc, int precision, int slices, void *stream);

/* ---- real GEMM without allocation ----
 * No counterpart in the reference: this entry is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * f4j_dgemm_dev with the same arguments and results, except that it never calls hipMalloc,
 * hipFree or an allocating routine: the modular form (SPLIT_INT8) keeps its residues,
 * exponents and check partials in a fixed 64 MiB device array of the library (one per GPU,
 * held until the process ends; about 1024^3 with 16 moduli fits), FP64 and STRICT need no
 * scratch, and the digit forms (SPLIT_BF16/FP16) return F4J_ERR_UNSUPPORTED.  A product whose
 * scratch would exceed the array returns F4J_ERR_INVALID_ARG.  Calls are serialised inside
 * the library and synchronous like f4j_dgemm_dev's split calls; the array keeps the last
 * call's contents.  Call f4j_init at set-up: the first call of the library initialises it. */
int f4j_dgemm_dev_noalloc(int opA, int opB, int64_t M, int64_t N, int64_t K,
                          const double *A, int64_t lda, const double *B, int64_t ldb,
                          double *C, int64_t ldc, int precision, int slices, void *stream);

/* ---- complex128 GEMM ----
 * Replaces, by (opA, opB):
 *   (N, N): CMatrix.mult(CMatrix)         M:arrays/dense/CMatrix.java:463-465 ->
 *           Cm128DeMatMultDispatcher       M:linalg/ops/dispatch/Cm128DeMatMultDispatcher.java:155-226 ->
 *           DenseSemiringMatMult.*         M:linalg/ops/dense/semiring_ops/DenseSemiringMatMult.java:54-424
 *   (N, T): CMatrix.multTranspose(CMatrix) M:arrays/backend/semiring_arrays/AbstractDenseSemiringMatrix.java:330-334 ->
 *           DenseSemiringMatMultTranspose.* M:linalg/ops/dense/semiring_ops/DenseSemiringMatMultTranspose.java:54-178
 *   (C, N): A.H().mult(B)                 M:arrays/backend/field_arrays/AbstractDenseFieldMatrix.java:98-102
 *           with the conjugate transpose fused into the operand load.
 * In the SPLIT precisions a complex product runs as three real split products (3M, Gauss)
 * when `use_3m` != 0, else four. */
int f4j_zgemm(int opA, int opB, int64_t M, int64_t N, int64_t K,
              const double *A, int64_t lda, const double *B, int64_t ldb,
              double *C, int64_t ldc, int precision, int slices, int use_3m);
int f4j_zgemm_dev(int opA, int opB, int64_t M, int64_t N, int64_t K,
                  const double *A, int64_t lda, const double *B, int64_t ldb,
                  double *C, int64_t ldc, int precision, int slices, int use_3m, void *stream);

/* ---- mixed real / complex products ----
 * Replace:
 *   Matrix.mult(CMatrix)  M:arrays/dense/Matrix.java:1793-1797, Matrix.mult(CVector) Matrix.java:1851-1853,
 *   CMatrix.mult(Matrix)  M:arrays/dense/CMatrix.java:878-881,  CMatrix.mult(Vector) CMatrix.java:894-896
 *   -> MatrixMultiplyDispatcher M:linalg/ops/MatrixMultiplyDispatcher.java:71-260
 *   -> RealFieldDenseMatMult.*  M:linalg/ops/dense/real_field_ops/RealFieldDenseMatMult.java:58-1090
 * f4j_dzgemm: C = A B with A real (M x K) and B complex (K x N); f4j_zdgemm: A complex (M x K) and
 * B real (K x N).  C is complex (M x N); no transposition; leading dimensions in each operand's
 * elements.  In the reference each component of C_ij is a k-ascending sum of
 * (component) x (real entry) products.  STRICT reproduces that bit for bit, as does AUTO for small
 * products (M*K + K*N < smallThreshold) and FP64 / AUTO for N = 1 (the matrix-vector forms).  FP64
 * otherwise runs the MFMA GEMM: f4j_dzgemm on B's interleaved storage viewed as a K x 2N real
 * matrix, f4j_zdgemm on [Re A; Im A] stacked into one 2M x K real matrix (the planes go through a
 * library-owned buffer, so f4j_zdgemm_dev synchronises `stream` before it returns).  The SPLIT
 * precisions run the split forms on the same real views. */
int f4j_dzgemm(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
               double *C, int64_t ldc, int precision, int slices);
int f4j_dzgemm_dev(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                   double *C, int64_t ldc, int precision, int slices, void *stream);
int f4j_zdgemm(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
               double *C, int64_t ldc, int precision, int slices);
int f4j_zdgemm_dev(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                   double *C, int64_t ldc, int precision, int slices, void *stream);

/* ---- matrix-vector products ----
 * Replace:
 *   real:    Matrix.mult(Vector)      M:arrays/dense/Matrix.java:1406-1408 ->
 *            ReDeMatVecMultDispatcher  M:linalg/ops/dispatch/ReDeMatVecMultDispatcher.java:141-181 ->
 *            RealDenseMatMult.standardVector / blockedVector / concurrentBlockedVector
 *                                      M:linalg/ops/dense/real/RealDenseMatMult.java:426-569
 *   complex: CMatrix.mult(CVector)    M:arrays/backend/semiring_arrays/AbstractDenseSemiringMatrix.java:295-298 ->
 *            DenseSemiringMatMultDispatcher.dispatchVector
 *                                      M:linalg/ops/dense/semiring_ops/DenseSemiringMatMultDispatcher.java:120-142 ->
 *            DenseSemiringMatMult.*Vector M:linalg/ops/dense/semiring_ops/DenseSemiringMatMult.java:435-574
 * y (M) = A (M x K, row-major, lda) x (K); incx / incy are the element strides of x and y (>= 1).
 * The results equal the reference's bit for bit: real products with M*K + K < smallThreshold are
 * k-ascending sums (standardVector); larger ones sum each 64-wide k block from +0.0 and add the
 * block sums in order (blockedVector); complex products are k-ascending sums of Complex128
 * products.  precision: F4J_PREC_AUTO, F4J_PREC_FP64 or F4J_PREC_STRICT, which all run these
 * kernels (the product is bound by reading A once); a SPLIT precision returns
 * F4J_ERR_UNSUPPORTED.  f4j_dgemm / f4j_zgemm with N = 1 and opA = N run the same kernels in
 * FP64 and STRICT precision (the reference's N = 1 dispatch, ReDeMatMultDispatcher.java:183-198). */
int f4j_dgemv(int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx,
              double *y, int64_t incy, int precision);
int f4j_dgemv_dev(int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx,
                  double *y, int64_t incy, int precision, void *stream);
int f4j_zgemv(int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx,
              double *y, int64_t incy, int precision);
int f4j_zgemv_dev(int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx,
                  double *y, int64_t incy, int precision, void *stream);

/* ---- Least squares: a caller of the product ----
 * Replaces RealLstsqSolver.solve(A, B) and solve(A, b)
 *   M:linalg/solvers/lstsq/LstsqSolver.java:157-181 (solve), :232-236 (decompose: Qh = Q.H(), R);
 *   M:linalg/decompositions/unitary/UnitaryDecomposition.java (decompose) and
 *   M:linalg/decompositions/unitary/RealUnitaryDecomposition.java (computeHouseholder,
 *   updateData, getQ :147-168): reduced Householder QR;
 *   M:linalg/transformations/Householder.java:156-188 (leftMultReflector);
 *   M:linalg/solvers/exact/triangular/RealBackSolver.java (solve; BackSolver.java:111-134 checks).
 * X (N x R, ldx) = argmin ||A X - B|| for A (M x N, lda) and B (M x R, ldb), M >= N >= 1; R = 1
 * is the Vector form.  The QR of A, Q and the back substitution run on the device in the
 * reference's order, bit for bit.  Y = Qh B is a library product at `precision`: with AUTO it
 * dispatches as f4j_dgemm does (R = 1 and the small path run the reference's order; larger
 * products run the configured default, FP64 MFMA), with STRICT the whole solve equals the
 * reference's bit for bit.  F4J_ERR_SHAPE: M < N or N = 0 (R must be square: ensureSquare).
 * F4J_ERR_SINGULAR: |det R| <= 2^-53 N or det R is NaN, det R being the product of R's diagonal
 * from the last entry up (checkSingular); X then holds the back substitution's values.  The _dev
 * form waits for `stream` before it returns (the check reads det R).  Library workspace: about
 * 3 M N + 2 N R doubles, grown and reused like the staging buffers. */
int f4j_dlstsq(int64_t M, int64_t N, int64_t R, const double *A, int64_t lda, const double *B, int64_t ldb,
               double *X, int64_t ldx, int precision);
int f4j_dlstsq_dev(int64_t M, int64_t N, int64_t R, const double *A, int64_t lda, const double *B, int64_t ldb,
                   double *X, int64_t ldx, int precision, void *stream);

/* ---- Pseudo-inverse: a caller of the product ----
 * Replaces Invert.pInv(Matrix) M:linalg/Invert.java:292-298 (V S^-1 U^T from RealSVD, S inverted by
 *   Invert.invDiag :126-145, which raises SingularMatrixException when |prod s_i| <= EPS_F64 k,
 *   EPS_F64 = 2^-53 (M:util/Flag4jConstants.java:52), k = min(M, N)).
 * X (N x M, ldx) = A^+ for A (M x N, lda), M, N >= 1.  For M >= N, A^+ = R^-1 Q^T from the reduced
 * Householder QR of A (the least-squares solve of f4j_dlstsq with B = I_M, whose product Q^T I is
 * exact and is skipped); for M < N, A^+ = ((A^T)^+)^T.  The pseudo-inverse of a full-rank matrix is
 * unique, so this equals the reference's V S^-1 U^T up to rounding; bit for bit it equals the
 * reference's own least-squares arithmetic (RealLstsqSolver.solve(A, I)).  F4J_ERR_SINGULAR when
 * |det R| <= 2^-53 min(M, N) or det R is NaN: |det R| = prod s_i, the reference's criterion; X then
 * holds the back substitution's values.  The _dev form waits for `stream` before it returns.
 * Library workspace: about 3 M N + 2 min(M, N) max(M, N) doubles, grown and reused. */
int f4j_dpinv(int64_t M, int64_t N, const double *A, int64_t lda, double *X, int64_t ldx);
int f4j_dpinv_dev(int64_t M, int64_t N, const double *A, int64_t lda, double *X, int64_t ldx, void *stream);

/* ---- grouped (batched) products ----
 * No reference counterpart: Flag4j multiplies one pair per call; a batch is a Java loop of
 * Matrix.mult (M:arrays/dense/Matrix.java:638-640) or A.H().mult(B) calls, each paying a
 * dispatcher lookup and a pool fork/join (M:concurrency/ThreadManager.java:155-183).
 * A plan fixes `count` independent products C_i (M_i x N_i) = op(A_i) op(B_i), all with the
 * same op codes, on DEVICE pointers (complex: interleaved, sizes in complex elements).  It
 * schedules their tiles over persistent launches (one per tile-shape bucket: each product goes
 * to the shape that pads it least per unit of efficiency; tile configuration per bucket from the
 * autotuner cache) and captures them into a hipGraph; f4j_grouped_plan_run replays the
 * graph on `stream` (asynchronous).  The operand pointers are baked into the plan: rerun it after
 * refilling the same buffers.  Precision: F4J_PREC_FP64 (or AUTO, which means FP64 here). */
typedef struct f4j_grouped_plan f4j_grouped_plan;
int f4j_grouped_plan_create(int is_complex, int opA, int opB, int64_t count,
                            const int64_t *M, const int64_t *N, const int64_t *K,
                            const double *const *A, const int64_t *lda,
                            const double *const *B, const int64_t *ldb,
                            double *const *C, const int64_t *ldc, int precision,
                            f4j_grouped_plan **plan);
/* Batched split-precision products (no reference counterpart: Flag4j is FP64-only; the batch
 * replaces a loop of Matrix.mult / A.H().mult(B) calls).  precision F4J_PREC_SPLIT_BF16 or
 * F4J_PREC_SPLIT_FP16 (slices = digits per operand, 0 = 7 (BF16) / 6 (FP16) for real products,
 * 8 / 7 for complex ones), or
 * F4J_PREC_SPLIT_INT8 (slices = moduli, 2..20; 0 = the fewest giving 64-bit operands at the
 * batch's largest K; inner sizes below 131072); complex products run as three real
 * products (3M, Gauss) when use3m != 0, else four.  splitCheckRows rows of every product are
 * verified after each run.  With splitFallback = 1 (at plan creation) the plan recomputes in
 * FP64, inside the same graph and without a host round trip, every product whose checked error
 * is above splitTolerance or whose operands hold NaN or Inf.  f4j_grouped_plan_create with a
 * SPLIT_* precision is this call with slices 0 and use3m 1.  Plans run, tune (a no-op: one
 * tile shape) and destroy like FP64 plans. */
int f4j_grouped_plan_create_split(int is_complex, int opA, int opB, int64_t count,
                                  const int64_t *M, const int64_t *N, const int64_t *K,
                                  const double *const *A, const int64_t *lda,
                                  const double *const *B, const int64_t *ldb,
                                  double *const *C, const int64_t *ldc, int precision, int slices,
                                  int use3m, f4j_grouped_plan **plan);
/* After a run of a split plan: err[p] = max |C - AB| / (|A||B|) of the split result over the
 * checked rows of problem p (-1 if none; NaN if problem p's operands hold NaN or Inf, which digits
 * and residues cannot carry), *products = real slice products per run.  Where err[p] is NaN or
 * above splitTolerance and the plan was created with splitFallback = 1, C_p holds the FP64
 * recomputation; otherwise C_p is the split result.  Waits for the last run. */
int f4j_grouped_plan_split_report(f4j_grouped_plan *plan, double *err, int64_t count, int *products);
int f4j_grouped_plan_run(f4j_grouped_plan *plan, void *stream);
int f4j_grouped_plan_destroy(f4j_grouped_plan *plan);
/* Work of the plan (flops, classical count) and, per shape bucket (nb <= 4 entries), the tile
 * configuration id and the number of tiles.  Ids 0-8: LDS-tiled kernels; 100-103: direct-load
 * kernels (one wave per 32x32, 64x32, 32x64 or 64x64 tile; real operands spanning < 2^31 bytes). */
int f4j_grouped_plan_info(const f4j_grouped_plan *plan, double *flops, int *tile_cfg, int64_t *tiles, int nb);
/* Autotuner: time every candidate tile configuration per bucket on this plan's own problems
 * (overwrites their C), keep the fastest in the plan and in the process-wide cache keyed by
 * (device arch, op kinds, bucket), and recapture the graph.  The cache persists via save/load
 * (text lines "arch kinds bucket id"; a direct-load id cached for a plan it cannot serve falls
 * back to the bucket's default).  Every configuration gives the same bits. */
int f4j_grouped_plan_tune(f4j_grouped_plan *plan, int reps);
int f4j_tuner_save(const char *path);
int f4j_tuner_load(const char *path);
int f4j_tuner_clear(void);

/* ---- split-precision report (no reference counterpart: Flag4j computes only in FP64) ----
 * What the calling thread's last SPLIT_* product delivered.  The FP64 residual check
 * recomputes `checked_rows` evenly spaced rows of C with compensated FP64 dot products on
 * the GPU and measures max |C - AB| / (|A||B|) over them. */
typedef struct f4j_split_report {
    int precision;        /* precision delivered: the SPLIT_* asked for, or FP64 after a fallback */
    int slices;           /* digit slices per operand */
    int products;         /* slice products computed (S (S + 1) / 2) */
    int launches;         /* slice-product kernel launches (the modular form: one for all real products) */
    int checked_rows;     /* rows verified by the residual check (0 = check off) */
    double max_rel_err;   /* max over the checked rows of |C - AB| / (|A||B|); -1 if not checked */
    double product_ms;    /* time of the slice-product launches, from HIP events on the call's stream,
                             when splitTiming = 1; -1 otherwise */
} f4j_split_report;
int f4j_last_split_report(f4j_split_report *out);

/* ---- dispatcher configuration ----
 * Mirrors M:linalg/ops/dispatch/configs/ReDeMatMultDispatchConfigs.java:52-117 and
 * Cm128DeMatMultDispatchConfigs.java:52-111 (key=value .properties files); unknown keys are
 * rejected with F4J_ERR_INVALID_ARG instead of being ignored.  Keys: smallThreshold,
 * defaultPrecision (fp64|strict|split_bf16|split_fp16|split_int8), defaultSlices,
 * tileConfig (-1 = per shape bucket, else a fixed FP64 tile id), splitCheckRows (rows of C the
 * FP64 residual check verifies after a split product; 0 = off), splitTolerance (largest accepted
 * |C - AB| / (|A||B|)), splitFallback (1 = recompute in FP64 when the check fails, 0 = return
 * F4J_ERR_ACCURACY), splitTiming (1 = report the slice-product time in f4j_split_report),
 * tunerCache (path of the grouped-GEMM tuning cache: read when the key is set and at f4j_init when
 * the file exists, rewritten by f4j_grouped_plan_tune; empty = none). */
int f4j_config_set(const char *key, const char *value);
int f4j_config_get(const char *key, char *buf, size_t buflen);
int f4j_config_load(const char *path);
/* The kernel the dispatcher would run for a real (is_complex = 0) or complex product of
 * op(A) (M x K) times op(B) (K x N): writes its name (e.g. "STRICT", "MFMA_F64_128x128x16"). */
int f4j_dispatch_name(int is_complex, int opA, int opB, int64_t M, int64_t N, int64_t K,
                      int precision, char *buf, size_t buflen);

#ifdef __cplusplus
}
#endif
#endif
