// Matrix-vector products in the reference's own summation orders, streamed
// from HBM once.
//
// Real, y = A x (A M x K row-major):
//   Matrix.mult(Vector) (M:arrays/dense/Matrix.java:1406-1408) ->
//   ReDeMatVecMultDispatcher (M:linalg/ops/dispatch/ReDeMatVecMultDispatcher.java:141-181), and
//   the N = 1 case of Matrix.mult(Matrix) (ReDeMatMultDispatcher.java:183-198) select
//   RealDenseMatMult.blockedVector / concurrentBlockedVector
//   (M:linalg/ops/dense/real/RealDenseMatMult.java:458-491, :537-569): y_i starts at +0.0 and,
//   for each 64-wide k block in order, y_i += (the block's k-ascending sum from +0.0).
//   Small problems take standardVector / standard (:426-447, :50-79), and
//   Matrix.multTranspose with a one-row B takes RealDenseMatMultTranspose (:53-82): both
//   are one k-ascending sum per row ("plain chain").
// Complex, y = A x:
//   CMatrix.mult(CVector) (M:arrays/backend/semiring_arrays/AbstractDenseSemiringMatrix.java:295-298
//   -> DenseSemiringMatMultDispatcher.dispatchVector :120-142 -> DenseSemiringMatMult.*Vector
//   :435-574) and the N = 1 case of CMatrix.mult(CMatrix): every variant reloads the running
//   sum from dest, so y_i is one k-ascending sum of Complex128.mult products
//   (M:numbers/Complex128.java:276-278, :336-341).
//
// Every result is a chain of dependent additions.  The plain real chain and
// complex A x form their products with the whole block and leave only the
// additions to one lane per chain (gemv_split_kernel, below).  The other
// kinds run their products inside the chains: a (row, 64-block) chain for the
// blocked form (8 rows x 8 blocks per wave, the block sums then added in block
// order across lanes) and a (row, component) pair for the mixed kinds (real A
// below about 32K rows takes 8 or 16 rows per wave, with the other lanes idle,
// to spread the rows over more waves).  The wave streams
// its slab of A in stages of 16 elements per chain (8 KB): coalesced 8-byte buffer loads into registers D stages ahead, then a
// write into LDS rotated so that the 64 lanes read their own chains' elements
// without bank conflicts.  Loads past K or past the last row are given an
// out-of-range offset and return zeros, which leave every chain unchanged (a
// sum that starts at +0.0 is never -0.0, so adding +0.0 is exact).  Products and
// sums are separately rounded (no contraction), so the results equal the
// reference's bit for bit.  The kernels are bound by reading A once.
#include <type_traits>

#include "amdgcn.h"
#include "f4j_internal.h"
#include "gemv_layout.h"

#pragma clang fp contract(off)

namespace f4j {
namespace {

constexpr uint32_t OOR = 0x80000000u;      // offset past every slab (slabs are < 2^31 bytes)

struct GemvArgs {
    int64_t M, K;
    const double *A;
    int64_t lda;              // doubles between rows (complex: 2 x complex lda)
    const double *x;
    int64_t incx;             // doubles between elements (complex: 2 x complex stride)
    double *y;
    int64_t incy;             // doubles between results (complex: 2 x complex stride)
    int64_t nstage;           // a multiple of the prefetch depth
    int conjx;
};

typedef double f64x2 __attribute__((ext_vector_type(2)));

__device__ inline __amdgpu_buffer_rsrc_t rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}

// Complex A (elements of two doubles) / complex x.
template <int KIND>
constexpr bool a_cplx() { return KIND == GEMV_CPLX || KIND == GEMV_CX; }
template <int KIND>
constexpr bool x_cplx() { return KIND == GEMV_CPLX || KIND == GEMV_RX; }

// Register-staged form of the blocked kind (A not 16-byte aligned, or odd K).
template <int KIND, int VEC, int D>
__global__ __launch_bounds__(64) void gemv_kernel(GemvArgs g) {
    static_assert(KIND == GEMV_BLOCKED, "the blocked form (other kinds: gemv_split_kernel)");
    constexpr int R = gemv::ROWS, XN = 2, NA = 16 / VEC;
    static_assert(D * (NA + XN) <= 63, "loads in flight must fit the vmcnt counter");
    __shared__ __attribute__((aligned(16))) double lds[1024 + 64 * XN];
    const int l = threadIdx.x;
    const int64_t row0 = (int64_t)blockIdx.x * R;
    const int rows = (int)((g.M - row0) < R ? (g.M - row0) : R);
    const __amdgpu_buffer_rsrc_t ra = rsrc(g.A + row0 * g.lda, (uint32_t)((int64_t)rows * g.lda * 8));
    const __amdgpu_buffer_rsrc_t rx = rsrc(g.x, (uint32_t)(((g.K - 1) * g.incx + 1) * 8));

    typedef typename std::conditional<VEC == 2, f64x2, double>::type AV;
    // Per-lane load geometry: k offset within a stage (a row past the slab
    // counts as past K) and byte offset at stage 0.  K < 2^28 (slabs < 2^31 bytes).
    const int K = (int)g.K, incx = (int)g.incx;
    int akl[NA], xkl[XN];
    uint32_t aoff[NA], xoff[XN];
    bool xneg[XN];
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        int r, kl;
        gemv::a_slot<VEC>(i, l, r, kl);
        akl[i] = r < rows ? kl : K;
        aoff[i] = (uint32_t)(((int64_t)r * g.lda + (int64_t)kl) * 8);
    }
#pragma unroll
    for (int h = 0; h < XN; ++h) {
        xkl[h] = gemv::x_slot(h, l);
        xneg[h] = false;
        xoff[h] = (uint32_t)((int64_t)xkl[h] * incx * 8);
    }
    AV va[D][NA];
    double vx[D][XN];
    auto issue = [&](AV (&a)[NA], double (&xv)[XN], int64_t s) {
        const int ks = gemv::kstage(s);
        const uint32_t ka = (uint32_t)ks * 8u, kx = (uint32_t)ks * (uint32_t)incx * 8u;
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const uint32_t off = ks + akl[i] < K ? aoff[i] + ka : OOR;
            if constexpr (VEC == 2)
                a[i] = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(ra, off, 0, 0));
            else
                a[i] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(ra, off, 0, 0));
        }
#pragma unroll
        for (int h = 0; h < XN; ++h) {
            const uint32_t off = ks + xkl[h] < K ? xoff[h] + kx : OOR;
            const double v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rx, off, 0, 0));
            xv[h] = xneg[h] ? -v : v;
        }
    };

    double s_acc = 0.0;    // this lane's chain
    double dest = 0.0;     // blocked form: the row's running sum (lanes l % 8 == 0)

#pragma unroll
    for (int d = 0; d < D; ++d) issue(va[d], vx[d], d);

    for (int64_t s0 = 0; s0 < g.nstage; s0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t s = s0 + d;
#pragma unroll
            for (int i = 0; i < NA; ++i) {
                if constexpr (VEC == 2) *(f64x2 *)&lds[gemv::a_store<2>(i, l)] = va[d][i];
                else lds[gemv::a_store<1>(i, l)] = va[d][i];
            }
#pragma unroll
            for (int h = 0; h < XN; ++h) lds[gemv::x_store(h, l)] = vx[d][h];
            issue(va[d], vx[d], s + D);       // past the last stage the offsets are out of range: zeros
            if constexpr (VEC == 2) {
#pragma unroll
                for (int p = 0; p < 8; ++p) {
                    const f64x2 a = *(const f64x2 *)&lds[gemv::a_read<2>(l, p)];
                    const double x0 = lds[gemv::x_read(l, 2 * p)];
                    const double x1 = lds[gemv::x_read(l, 2 * p + 1)];
                    s_acc = __dadd_rn(s_acc, __dmul_rn(a.x, x0));
                    s_acc = __dadd_rn(s_acc, __dmul_rn(a.y, x1));
                }
            } else {
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const double a = lds[gemv::a_read<1>(l, e)];
                    const double xv = lds[gemv::x_read(l, e)];
                    s_acc = __dadd_rn(s_acc, __dmul_rn(a, xv));
                }
            }
            if ((s & 3) == 3) {
                // the 8 block sums of row l / 8 sit in lanes 8 (l / 8) + j: add them in block order
                double t = dest;
#pragma unroll
                for (int j = 0; j < 8; ++j) t = __dadd_rn(t, __shfl(s_acc, (l & ~7) + j));
                dest = t;
                s_acc = 0.0;
            }
        }
    }

    const int r = l >> 3;
    if ((l & 7) == 0 && r < rows) g.y[(row0 + r) * g.incy] = dest;
}

// ---- direct-to-LDS form (16-byte aligned A with even K and lda; complex always) ----
//
// The same chains and LDS layout as above, but every load writes LDS directly
// (buffer_load ... lds), into a ring of D stage slots, D - 1 stages ahead.
// Loads produce no registers, so the waits are explicit: each step issues the
// same number of loads (offsets clamped into the slab), and at the top of step
// s vmcnt((D - 2) L) retires stage s.  Elements past K are skipped by the
// reader (a uniform test picks the masked path for the last stages only).
//
// It serves the blocked form (the other kinds run gemv_split_kernel): NL = 8
// A loads per stage, 8 rows x 8 blocks of chains per wave.

template <int KIND, int XV, int D, int NL>
__global__ __launch_bounds__(64) void gemv_dma_kernel(GemvArgs g) {
    static_assert(KIND == GEMV_BLOCKED && NL == gemv::DMA_NL, "the blocked form (other kinds: gemv_split_kernel)");
    constexpr int R = gemv::ROWS, XL = gemv::x_dma_loads<XV>(), L = gemv::dma_loads<XV>();
    constexpr int SLOT = gemv::slot_bytes();
    constexpr int ew = a_cplx<KIND>() ? 2 : 1, xw = x_cplx<KIND>() ? 2 : 1;
    static_assert((D - 1) * L <= 63, "loads in flight must fit the vmcnt counter");
    __shared__ __attribute__((aligned(1024))) char lds[D * SLOT];
    const int l = threadIdx.x;
    const int64_t row0 = (int64_t)blockIdx.x * R;
    const int rows = (int)((g.M - row0) < R ? (g.M - row0) : R);
    const int K = (int)g.K, incx = (int)g.incx;
    const __amdgpu_buffer_rsrc_t ra = rsrc(g.A + row0 * g.lda, (uint32_t)((int64_t)rows * g.lda * 8));
    const __amdgpu_buffer_rsrc_t rx = rsrc(g.x, (uint32_t)(((g.K - 1) * g.incx + xw) * 8));
    // last valid load start: a pair (real) or an element (complex) / x element
    const int ka_max = a_cplx<KIND>() ? K - 1 : K - 2;
    const int kx_max = XV == 2 ? K - 2 : K - 1;

    // per-lane load geometry
    int akl[8];           // (sized 8, not NL: an array of template-dependent size captured by the
    uint32_t arow[8];     //  lambdas below makes hipcc drop the kernel's host stub without a diagnostic)
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        int r, kl;
        gemv::a_slot<2>(i, l, r, kl);
        akl[i] = kl;
        arow[i] = (uint32_t)((int64_t)(r < rows ? r : rows - 1) * g.lda * 8);
    }
    int xkl[XL], xadd[XL];
#pragma unroll
    for (int h = 0; h < XL; ++h) gemv::dma_x_slot<XV>(h, l, xkl[h], xadd[h]);

    // The LDS destination of a load is the address given for lane 0; lane l's
    // bytes follow at l times the load size (gemv::dma_a_store, dma_x_store).
    auto issue = [&](int64_t s) {
        const int ks = gemv::kstage(s);
        char *slot = lds + (int)(s % D) * SLOT;
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int k = min(ks + akl[i], ka_max);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_void *)(slot + gemv::dma_a_store(i, 0)), 16,
                                                     arow[i] + (uint32_t)k * (8u * ew), 0, 0, 0);
        }
#pragma unroll
        for (int h = 0; h < XL; ++h) {
            const int k = min(ks + xkl[h], kx_max);
            const uint32_t off = (uint32_t)k * (uint32_t)incx * 8u + xadd[h];   // incx in doubles
            if constexpr (XV == 2)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rx, (lds_void *)(slot + gemv::dma_x_store<2>(h, 0)), 16, off,
                                                         0, 0, 0);
            else
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rx, (lds_void *)(slot + gemv::dma_x_store<1>(h, 0)), 4, off,
                                                         0, 0, 0);
        }
    };

    // lane l runs chain (row l / 8, block l % 8): k = kstage + kb + 2 p, 2 p + 1
    const int kb = 64 * (l & 7);

    double s_acc = 0.0, dest = 0.0;
    auto compute = [&](int64_t s, bool masked) {
        const char *slot = lds + (int)(s % D) * SLOT;
        const int ks = gemv::kstage(s);
        {
#pragma unroll
            for (int p = 0; p < 8; ++p) {
                const f64x2 a = *(const f64x2 *)(slot + gemv::dma_a_read(l, p));
                double x0, x1;
                if constexpr (XV == 2) {
                    const f64x2 w = *(const f64x2 *)(slot + gemv::dma_x_read<2>(l, p, 0));
                    x0 = w.x;
                    x1 = w.y;
                } else {
                    x0 = *(const double *)(slot + gemv::dma_x_read<1>(l, p, 0));
                    x1 = *(const double *)(slot + gemv::dma_x_read<1>(l, p, 1));
                }
                const double v = __dadd_rn(__dadd_rn(s_acc, __dmul_rn(a.x, x0)), __dmul_rn(a.y, x1));
                s_acc = (!masked || ks + kb + 2 * p < K) ? v : s_acc;
            }
        }
    };

#pragma unroll
    for (int d = 0; d < D - 1; ++d) issue(d);
    for (int64_t s = 0; s < g.nstage; ++s) {
        __builtin_amdgcn_sched_barrier(0);
        F4J_WAIT_VMCNT(gemv::dma_wait(D, L));
        __builtin_amdgcn_sched_barrier(0);
        issue(s + D - 1);                 // past the last stage: clamped reloads, never read
        __builtin_amdgcn_sched_barrier(0);
        if (gemv::kstage(s) + gemv::SPAN <= K) compute(s, false);
        else compute(s, true);
        if ((s & 3) == 3) {                // the row's 8 block sums, added in block order
            double t = dest;
#pragma unroll
            for (int b = 0; b < 8; ++b) t = __dadd_rn(t, __shfl(s_acc, (l & ~7) + b));
            dest = t;
            s_acc = 0.0;
        }
    }
    F4J_WAIT_VMCNT(0);               // no LDS write in flight when the wave ends

    const int r = l >> 3;
    if ((l & 7) == 0 && r < rows) g.y[(row0 + r) * g.incy] = dest;
}

template <int KIND, int XV, int D, int NL = gemv::DMA_NL>
void launch_dma(GemvArgs g, int64_t stages, hipStream_t stream) {
    constexpr int R = gemv::ROWS;
    g.nstage = stages;
    hipLaunchKernelGGL((gemv_dma_kernel<KIND, XV, D, NL>), dim3((unsigned)((g.M + R - 1) / R)), dim3(64), 0, stream,
                       g);
}

// One-chain-per-row products with the products spread over the block and only
// the additions in chains: complex A x (GEMV_CPLX: Complex128.mult products,
// one chain per row and component), the mixed kinds (GEMV_RX: x_k.re a_ik and
// x_k.im a_ik; GEMV_CX: a_ik.re x_k and a_ik.im x_k, per component) and the
// real plain chain (GEMV_CHAIN).  A
// chain's serial work is its additions: the products of one k chunk are
// independent.  A block takes R rows; every chunk of KC = 64 k, all 256
// threads form the chunk's R x 64 products (E = R / 4 consecutive k of one row
// per thread; complex re = ar xr - ai xi, im = ar xi + ai xr; every product
// and sum rounded, as the reference's) into an LDS buffer, while the first C R
// lanes of wave 0 (row l % R, component l / R) add the previous chunk's
// products in k order.  The buffers alternate, one barrier per chunk, and
// every thread's loads for a chunk are issued one chunk ahead (complex, M = K =
// 8192: 205 us; 8 or 16 rows per block alike; 128-k chunks, or loads two
// chunks ahead, 252 us; real chain: 8192^2 150 us, 32768 x 2048 111 us, with
// 128-k chunks 152 / 168 us).  With one lane per chain (complex: 32 rows per
// wave, one wave per CU at M = 8K) the chains' own instruction latency bound
// it (complex 343 us, real chain 8192^2 191 us).
template <int KIND, int R>
__global__ __launch_bounds__(256) void gemv_split_kernel(GemvArgs g) {
    static_assert(KIND != GEMV_BLOCKED, "one chain per row and component");
    constexpr bool CPLX = KIND == GEMV_CPLX, AC = a_cplx<KIND>(), XC = x_cplx<KIND>();
    constexpr int C = (AC || XC) ? 2 : 1;                     // components of y
    constexpr int KC = 64, E = KC * R / 256, TPR = KC / E;   // k per chunk; products per thread; threads per row
    static_assert(E * 256 == KC * R && E >= 1 && E <= 8, "whole products per thread");
    typedef typename std::conditional<AC, f64x2, double>::type VA;
    typedef typename std::conditional<XC, f64x2, double>::type VX;
    typedef typename std::conditional<C == 2, f64x2, double>::type V;
    __shared__ V prod[2][R][KC + 1];             // +1: the chain lanes read distinct banks
    const int t = threadIdx.x;
    const int r = t / TPR, q = t % TPR;
    const int64_t row0 = (int64_t)blockIdx.x * R;
    const int64_t row = row0 + r;
    const bool row_ok = row < g.M;
    const double *arow = g.A + (row_ok ? row : 0) * g.lda;
    const int64_t nchunk = (g.K + KC - 1) / KC;
    const double xs = g.conjx ? -1.0 : 1.0;      // conj(x): the imaginary part's sign flipped exactly

    VA a[E];
    VX x[E];
    auto load = [&](int64_t c) {                 // chunk c's A and x elements of this thread
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const int64_t k = c * KC + E * q + u;
            const bool ok = row_ok && k < g.K;
            a[u] = ok ? *(const VA *)(arow + (AC ? 2 : 1) * k) : VA{};
            x[u] = k < g.K ? *(const VX *)(g.x + k * g.incx) : VX{};
        }
    };
    auto products = [&](int buf) {
#pragma unroll
        for (int u = 0; u < E; ++u) {
            if constexpr (CPLX) {
                const double xr = x[u].x, xi = xs * x[u].y;
                const double re = __dsub_rn(__dmul_rn(a[u].x, xr), __dmul_rn(a[u].y, xi));
                const double im = __dadd_rn(__dmul_rn(a[u].x, xi), __dmul_rn(a[u].y, xr));
                prod[buf][r][E * q + u] = f64x2{re, im};
            } else if constexpr (KIND == GEMV_RX) {     // x.mult(a) = (x.re a, x.im a)
                prod[buf][r][E * q + u] = f64x2{__dmul_rn(x[u].x, a[u]), __dmul_rn(x[u].y, a[u])};
            } else if constexpr (KIND == GEMV_CX) {     // a.mult(x) = (a.re x, a.im x)
                prod[buf][r][E * q + u] = f64x2{__dmul_rn(a[u].x, x[u]), __dmul_rn(a[u].y, x[u])};
            } else {
                prod[buf][r][E * q + u] = __dmul_rn(a[u], x[u]);
            }
        }
    };
    // chain lanes: wave 0, lanes 0 .. C R - 1
    const bool chain = t < C * R;
    const int cr = t % R, comp = t / R;
    double s = 0.0;

    load(0);
    products(0);
    if (nchunk > 1) load(1);
    __syncthreads();
    for (int64_t c = 0; c < nchunk; ++c) {
        if (c + 1 < nchunk) {
            products((int)((c + 1) & 1));        // chunk c + 1, loaded one chunk ago
            if (c + 2 < nchunk) load(c + 2);
        }
        if (chain) {
            const int kn = (int)min<int64_t>(KC, g.K - c * KC);
            const double *pc = (const double *)&prod[c & 1][cr][0] + comp;
            if (kn == KC) {                      // full chunk: unrolled, the reads run ahead of the adds
#pragma unroll
                for (int k = 0; k < KC; ++k) s = __dadd_rn(s, pc[C * k]);
            } else {
                for (int k = 0; k < kn; ++k) s = __dadd_rn(s, pc[C * k]);
            }
        }
        __syncthreads();
    }
    if (chain && row0 + cr < g.M) g.y[(row0 + cr) * g.incy + comp] = s;
}

template <int KIND, int VEC, int D>
void launch(GemvArgs g, int64_t stages, hipStream_t stream) {
    constexpr int R = gemv::ROWS;
    g.nstage = (stages + D - 1) / D * D;
    hipLaunchKernelGGL((gemv_kernel<KIND, VEC, D>), dim3((unsigned)((g.M + R - 1) / R)), dim3(64), 0, stream, g);
}

}  // namespace

namespace {
int a_width(int kind) { return kind == GEMV_CPLX || kind == GEMV_CX ? 2 : 1; }
int x_width(int kind) { return kind == GEMV_CPLX || kind == GEMV_RX ? 2 : 1; }
int y_width(int kind) { return kind == GEMV_BLOCKED || kind == GEMV_CHAIN ? 1 : 2; }
int rows_of(int kind) { return kind == GEMV_BLOCKED ? 8 : (a_width(kind) == 2 ? 32 : 64); }
}  // namespace

bool gemv_supported(int kind, int64_t K, int64_t lda, int64_t incx) {
    // 32-bit offsets: the wave's slab of A and all of x stay below 2^31 bytes,
    // also for the stage offsets computed past the end (K + 4096 elements).
    const int64_t lim = 0x7FFFFFF0;
    return K >= 1 && lda >= K && incx >= 1 && rows_of(kind) * lda * a_width(kind) * 8 <= lim &&
           (K + 4096) * incx * x_width(kind) * 8 <= lim;
}

bool gemv_applicable(int kind, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx) {
    if (!gemv_supported(kind, K, lda, incx)) return false;
    // The blocked form picks its load path from the alignment itself.  The split
    // kernel reads real entries as 8-byte doubles and complex ones as 16-byte
    // pairs: complex A and complex x must be 16-byte aligned.
    if (kind == GEMV_BLOCKED) return true;
    const bool a_ok = a_width(kind) == 1 || ((uintptr_t)A & 15) == 0;
    const bool x_ok = x_width(kind) == 1 || ((uintptr_t)x & 15) == 0;
    return a_ok && x_ok;
}

hipError_t launch_gemv(int kind, int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx,
                       double *y, int64_t incy, bool conjx, hipStream_t stream) {
    if (M == 0) return hipSuccess;
    if (!gemv_applicable(kind, K, A, lda, x, incx)) return hipErrorInvalidValue;
    GemvArgs g{M, K, A, lda * a_width(kind), x, incx * x_width(kind), y, incy * y_width(kind), 0, conjx ? 1 : 0};
    // Direct-to-LDS loads move 16 bytes: 16-byte aligned pairs (complex: elements)
    // that never straddle a row end.  Otherwise the register-staged form.
    const bool a16 = ((uintptr_t)A & 15) == 0 && (a_width(kind) == 2 || (lda % 2 == 0 && K % 2 == 0));
    const bool x16 = ((uintptr_t)x & 15) == 0 && incx == 1 && K % 2 == 0;
    if (kind == GEMV_BLOCKED) {
        const int64_t stages = gemv::stages_for(K);
        if (a16 && x16) launch_dma<GEMV_BLOCKED, 2, 4>(g, stages, stream);
        else if (a16) launch_dma<GEMV_BLOCKED, 1, 4>(g, stages, stream);
        else launch<GEMV_BLOCKED, 1, 3>(g, stages, stream);
    } else if (kind == GEMV_CHAIN) {
        constexpr int R = 16;
        hipLaunchKernelGGL((gemv_split_kernel<GEMV_CHAIN, R>), dim3((unsigned)((M + R - 1) / R)), dim3(256), 0,
                           stream, g);
    } else if (kind == GEMV_RX) {
        constexpr int R = 16;
        hipLaunchKernelGGL((gemv_split_kernel<GEMV_RX, R>), dim3((unsigned)((M + R - 1) / R)), dim3(256), 0, stream,
                           g);
    } else if (kind == GEMV_CX) {
        constexpr int R = 16;
        hipLaunchKernelGGL((gemv_split_kernel<GEMV_CX, R>), dim3((unsigned)((M + R - 1) / R)), dim3(256), 0, stream,
                           g);
    } else {
        constexpr int R = 16;
        hipLaunchKernelGGL((gemv_split_kernel<GEMV_CPLX, R>), dim3((unsigned)((M + R - 1) / R)), dim3(256), 0,
                           stream, g);
    }
    return hipGetLastError();
}

}  // namespace f4j
