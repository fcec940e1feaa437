// C ABI of flag4j_amd: lifecycle, device buffers, dispatcher and GEMM entries.
//
// The dispatcher here stands where Flag4j's ReDeMatMultDispatcher /
// Cm128DeMatMultDispatcher stand (M:linalg/ops/dispatch/ReDeMatMultDispatcher.java:157-237,
// Cm128DeMatMultDispatcher.java:155-226; M: = src/main/java/org/flag4j/): it
// takes the shapes of a product and picks the kernel.  Products whose operands
// hold fewer than `smallThreshold` entries in total (the reference's small
// path, ReDeMatMultDispatcher.java:159-168) run the bitwise-strict kernel;
// larger ones run the configured precision on the matrix cores with a tile
// shape chosen per shape bucket.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <atomic>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/flag4j_amd.h"
#include "f4j_internal.h"
#include "gemm_tile.h"
#include "crt_gemm.h"
#include "split_batch.h"
#include "split_gemm.h"

namespace f4j {

namespace {

thread_local std::string t_err;
thread_local f4j_split_report t_split{};

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct Config {
    int64_t small_threshold = 2450;   // ReDeMatMultDispatchConfigs.java:63 default
    int default_precision = F4J_PREC_FP64;
    int default_slices = 0;
    int tile_config = -1;             // -1: chosen per shape bucket; else a fixed FP64 tile id
    int split_check_rows = 4;         // rows of C checked in FP64 after a split product (0 = off)
    double split_tolerance = 1e-15;   // max |C - AB| / (|A||B|) accepted on the checked rows
    int split_fallback = 1;           // 1: recompute in FP64 when the check fails; 0: return F4J_ERR_ACCURACY
    int split_timing = 0;             // 1: time the slice-product launches with HIP events (report product_ms)
    std::string tuner_cache;          // grouped-GEMM tuning cache file: read at init, written after tuning
};

// Scratch memory of the allocation-free entry (f4j_dgemm_dev_noalloc): the
// modular form's residues, exponents and check partials live here instead of in
// the grown workspace.  Each GPU the library runs on holds its own copy until the
// process ends.  In host code `scratch` names no device address: the address of
// a device's copy comes from hipGetSymbolAddress at initialisation
// (State::scratch_ptr), and only that address is handed to kernels.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Staging of the host-array calls.  The caller's (pageable) arrays move through
// a ring of pinned chunks: the calling thread copies chunk i + 1 in while the DMA
// engine moves chunk i, and on the way out the DMA of the next chunks runs while
// it copies one out (pinned DMA reaches about 56 GB/s each way on MI355X hosts,
// a pageable hipMemcpy 18-30 GB/s).  Each context owns its device buffers,
// stream and ring, so up to STAGE_CTX host calls run at once; further callers
// wait for a context.
constexpr int STAGE_CTX = 4;
constexpr int STAGE_SLOTS = 4;
constexpr size_t STAGE_SLOT_BYTES = size_t(16) << 20;
constexpr int STAGE_COPY_THREADS = 4;      // host copy of one chunk split over this many threads

struct Stage {
    std::mutex mu;                         // held by the call using the context
    bool ready = false;
    hipStream_t st = nullptr;              // copies and products of this context, in order
    DevBuf A, B, C;
    void *slot[STAGE_SLOTS] = {};          // pinned (hipHostMalloc)
    hipEvent_t slot_ev[STAGE_SLOTS] = {};  // last DMA that used the slot
    bool slot_used[STAGE_SLOTS] = {};
};

struct State {
    std::mutex mu;                    // guards init state
    std::mutex cfg_mu;                // guards cfg
    bool inited = false;
    int device = -1;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    Config cfg;
    Stage stages[STAGE_CTX];
    std::mutex ws_mu;                 // guards ws and tev (split-precision workspace, product timing)
    DevBuf ws;
    std::mutex mix_mu;                // guards mixws (planes of complex x real products)
    DevBuf mixws;
    std::mutex ls_mu;                 // guards lsws (least-squares factors and products)
    DevBuf lsws;
    hipEvent_t tev[8] = {};                 // product timing: a begin / end pair per real product (up to 4)
    hipStream_t side = nullptr;             // B's split and the residual sums beside A's split and the
                                            // modular products (guarded by ws_mu)
    hipEvent_t fork_ev = nullptr, join_ev = nullptr, split_fork_ev = nullptr, split_join_ev = nullptr;
    void *scratch_ptr[NOALLOC_MAX_DEVICES] = {};   // per device: its copy of `scratch` (set at init)
    // The last complex digit-form product's batch plan (its tables, digit pools and
    // schedule are fixed by the arguments below), reused by a call with the same
    // ones instead of being built and freed again (guarded by ws_mu).
    struct CplxSplitKey {
        int opA, opB, carrier, slices, use3m, check_rows;
        int64_t M, N, K, lda, ldb, ldc;
        const double *A, *B;
        double *C;
        bool operator==(const CplxSplitKey &o) const {
            return opA == o.opA && opB == o.opB && carrier == o.carrier && slices == o.slices && use3m == o.use3m &&
                   check_rows == o.check_rows && M == o.M && N == o.N && K == o.K && lda == o.lda && ldb == o.ldb &&
                   ldc == o.ldc && A == o.A && B == o.B && C == o.C;
        }
    };
    SplitBatch *cplx_split = nullptr;
    CplxSplitKey cplx_split_key{};
    // Pinned host words the split forms read back (non-finite flag, checked error),
    // so those copies do not stall the host between launches (guarded by ws_mu).
    struct Readback {
        int nonfinite;
        double err;
    };
    Readback *rb = nullptr;
};

State &S() {
    static State s;
    return s;
}

int fail(int code, const std::string &msg) {
    t_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? F4J_ERR_OUT_OF_MEMORY : F4J_ERR_HIP;
    return fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

#define F4J_HIP(call, what)                        \
    do {                                           \
        hipError_t e_ = (call);                    \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

int init_locked(State &s, int device) {
    if (s.inited) {
        if (device == s.device) return F4J_OK;
        return fail(F4J_ERR_INVALID_ARG, "f4j_init: already initialised on device " + std::to_string(s.device) +
                                             "; call f4j_shutdown first");
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(F4J_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n)
        return fail(F4J_ERR_INVALID_ARG, "device " + std::to_string(device) + " out of range [0," + std::to_string(n) + ")");
    F4J_HIP(hipSetDevice(device), "hipSetDevice");
    F4J_HIP(hipGetDeviceProperties(&s.prop, device), "hipGetDeviceProperties");
    if (device < NOALLOC_MAX_DEVICES && !s.scratch_ptr[device]) {
        void *p = nullptr;        // this device's copy of the scratch array (the device is current)
        F4J_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress (allocation-free scratch)");
        s.scratch_ptr[device] = p;
    }
    F4J_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
    for (hipEvent_t &e : s.tev) F4J_HIP(hipEventCreate(&e), "hipEventCreate");
    F4J_HIP(hipStreamCreateWithFlags(&s.side, hipStreamNonBlocking), "hipStreamCreate");
    F4J_HIP(hipEventCreateWithFlags(&s.fork_ev, hipEventDisableTiming), "hipEventCreate");
    F4J_HIP(hipEventCreateWithFlags(&s.join_ev, hipEventDisableTiming), "hipEventCreate");
    F4J_HIP(hipEventCreateWithFlags(&s.split_fork_ev, hipEventDisableTiming), "hipEventCreate");
    F4J_HIP(hipEventCreateWithFlags(&s.split_join_ev, hipEventDisableTiming), "hipEventCreate");
    if (!s.rb) F4J_HIP(hipHostMalloc((void **)&s.rb, sizeof(State::Readback), hipHostMallocDefault), "hipHostMalloc");
    s.device = device;
    s.inited = true;
    return F4J_OK;
}

// Load the configured tuning cache, if the file exists (a missing file is the
// first run; a malformed one leaves the cache as it was).
void load_tuner_cache() {
    std::string path;
    {
        std::lock_guard<std::mutex> g(S().cfg_mu);
        path = S().cfg.tuner_cache;
    }
    if (path.empty()) return;
    std::ifstream probe(path);
    if (!probe) return;
    const std::string keep = t_err;
    if (f4j_tuner_load(path.c_str()) != F4J_OK) t_err = keep;
}

// Make sure the library is initialised and the calling thread's current
// device is the library's device.
int ensure_ready() {
    State &s = S();
    {
        std::lock_guard<std::mutex> g(s.mu);
        if (!s.inited) {
            const int rc = init_locked(s, 0);
            if (rc != F4J_OK) return rc;
            load_tuner_cache();
        }
    }
    F4J_HIP(hipSetDevice(s.device), "hipSetDevice");
    return F4J_OK;
}

hipStream_t pick_stream(void *stream) { return stream ? (hipStream_t)stream : S().stream; }

int grow(DevBuf &b, size_t bytes) {
    if (bytes <= b.cap) return F4J_OK;
    size_t cap = std::max<size_t>(bytes, b.cap + b.cap / 2);
    cap = (cap + 4095) & ~size_t(4095);
    if (b.p) {
        F4J_HIP(hipFree(b.p), "hipFree");
        b.p = nullptr;
        b.cap = 0;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, cap);
    if (e != hipSuccess && cap > bytes) {
        (void)hipGetLastError();
        cap = bytes;
        e = hipMalloc(&p, cap);
    }
    if (e != hipSuccess) return hip_fail(e, "hipMalloc (staging buffer)");
    b.p = p;
    b.cap = cap;
    return F4J_OK;
}

// ---- host staging ----

int stage_setup(Stage &g) {
    if (g.ready) return F4J_OK;
    F4J_HIP(hipStreamCreateWithFlags(&g.st, hipStreamNonBlocking), "hipStreamCreate (staging)");
    for (int i = 0; i < STAGE_SLOTS; ++i) {
        F4J_HIP(hipHostMalloc(&g.slot[i], STAGE_SLOT_BYTES, hipHostMallocDefault), "hipHostMalloc (staging ring)");
        F4J_HIP(hipEventCreateWithFlags(&g.slot_ev[i], hipEventDisableTiming), "hipEventCreate (staging)");
        g.slot_used[i] = false;
    }
    g.ready = true;
    return F4J_OK;
}

void stage_release(Stage &g) {
    if (!g.ready) return;
    (void)hipStreamSynchronize(g.st);
    for (int i = 0; i < STAGE_SLOTS; ++i) {
        if (g.slot[i]) (void)hipHostFree(g.slot[i]);
        if (g.slot_ev[i]) (void)hipEventDestroy(g.slot_ev[i]);
        g.slot[i] = nullptr;
        g.slot_ev[i] = nullptr;
    }
    for (DevBuf *b : {&g.A, &g.B, &g.C})
        if (b->p) {
            (void)hipFree(b->p);
            b->p = nullptr;
            b->cap = 0;
        }
    (void)hipStreamDestroy(g.st);
    g.st = nullptr;
    g.ready = false;
}

// A free context if there is one, else wait for one (round robin).
Stage &stage_acquire(std::unique_lock<std::mutex> &lk) {
    State &s = S();
    for (Stage &g : s.stages) {
        std::unique_lock<std::mutex> l(g.mu, std::try_to_lock);
        if (l.owns_lock()) {
            lk = std::move(l);
            return g;
        }
    }
    static std::atomic<unsigned> next{0};
    Stage &g = s.stages[next++ % STAGE_CTX];
    lk = std::unique_lock<std::mutex>(g.mu);
    return g;
}

// Copy `rows` rows of `row_bytes` between a strided host array (pitch bytes) and
// a packed buffer, on up to STAGE_COPY_THREADS threads for large chunks.
void host_rows_copy(char *packed, const char *host_src, char *host_dst, int64_t rows, size_t row_bytes, size_t pitch) {
    auto part = [&](int64_t r0, int64_t r1) {
        if (host_src) {
            if (pitch == row_bytes) std::memcpy(packed + r0 * row_bytes, host_src + r0 * pitch, (r1 - r0) * row_bytes);
            else
                for (int64_t r = r0; r < r1; ++r) std::memcpy(packed + r * row_bytes, host_src + r * pitch, row_bytes);
        } else {
            if (pitch == row_bytes) std::memcpy(host_dst + r0 * pitch, packed + r0 * row_bytes, (r1 - r0) * row_bytes);
            else
                for (int64_t r = r0; r < r1; ++r) std::memcpy(host_dst + r * pitch, packed + r * row_bytes, row_bytes);
        }
    };
    const size_t bytes = (size_t)rows * row_bytes;
    const int nt = (int)std::min<int64_t>(rows, bytes >= (size_t(4) << 20) ? STAGE_COPY_THREADS : 1);
    if (nt <= 1) {
        part(0, rows);
        return;
    }
    std::vector<std::thread> th;
    const int64_t per = (rows + nt - 1) / nt;
    for (int t = 1; t < nt; ++t) {
        const int64_t r0 = std::min(rows, t * per), r1 = std::min(rows, (t + 1) * per);
        if (r0 < r1) th.emplace_back(part, r0, r1);
    }
    part(0, std::min(rows, per));
    for (std::thread &x : th) x.join();
}

// Host (pageable, `pitch` bytes between rows) -> device (packed rows), through the ring.
int stage_in(Stage &g, void *dev, const void *host, int64_t rows, size_t row_bytes, size_t pitch) {
    if (rows <= 0 || row_bytes == 0) return F4J_OK;
    if (row_bytes > STAGE_SLOT_BYTES) {       // rows longer than a chunk: the runtime's own path
        F4J_HIP(hipMemcpy2DAsync(dev, row_bytes, host, pitch, row_bytes, rows, hipMemcpyHostToDevice, g.st),
                "copy to device");
        return F4J_OK;
    }
    const int64_t rpc = (int64_t)(STAGE_SLOT_BYTES / row_bytes);
    int k = 0;
    for (int64_t r0 = 0; r0 < rows; r0 += rpc, ++k) {
        const int64_t n = std::min(rpc, rows - r0);
        const int sl = k % STAGE_SLOTS;
        if (g.slot_used[sl]) F4J_HIP(hipEventSynchronize(g.slot_ev[sl]), "staging wait");
        host_rows_copy((char *)g.slot[sl], (const char *)host + r0 * pitch, nullptr, n, row_bytes, pitch);
        F4J_HIP(hipMemcpyAsync((char *)dev + r0 * row_bytes, g.slot[sl], n * row_bytes, hipMemcpyHostToDevice, g.st),
                "copy to device");
        F4J_HIP(hipEventRecord(g.slot_ev[sl], g.st), "hipEventRecord (staging)");
        g.slot_used[sl] = true;
    }
    return F4J_OK;
}

// Device (packed rows) -> host (pageable, `pitch` bytes between rows), through the
// ring; waits for the work queued before it on the context's stream.  Synchronous.
int stage_out(Stage &g, void *host, size_t pitch, const void *dev, int64_t rows, size_t row_bytes) {
    if (rows <= 0 || row_bytes == 0) return F4J_OK;
    if (row_bytes > STAGE_SLOT_BYTES) {
        F4J_HIP(hipMemcpy2DAsync(host, pitch, dev, row_bytes, row_bytes, rows, hipMemcpyDeviceToHost, g.st),
                "copy to host");
        F4J_HIP(hipStreamSynchronize(g.st), "stream synchronize");
        return F4J_OK;
    }
    const int64_t rpc = (int64_t)(STAGE_SLOT_BYTES / row_bytes);
    const int64_t nchunk = (rows + rpc - 1) / rpc;
    auto issue = [&](int64_t k) -> int {
        const int sl = (int)(k % STAGE_SLOTS);
        const int64_t r0 = k * rpc, n = std::min(rpc, rows - r0);
        F4J_HIP(hipMemcpyAsync(g.slot[sl], (const char *)dev + r0 * row_bytes, n * row_bytes, hipMemcpyDeviceToHost,
                               g.st),
                "copy to host");
        F4J_HIP(hipEventRecord(g.slot_ev[sl], g.st), "hipEventRecord (staging)");
        g.slot_used[sl] = true;
        return F4J_OK;
    };
    // keep STAGE_SLOTS - 1 chunks in flight ahead of the one being copied out
    for (int64_t k = 0; k < std::min<int64_t>(nchunk, STAGE_SLOTS - 1); ++k)
        if (const int rc = issue(k); rc != F4J_OK) return rc;
    for (int64_t k = 0; k < nchunk; ++k) {
        const int sl = (int)(k % STAGE_SLOTS);
        F4J_HIP(hipEventSynchronize(g.slot_ev[sl]), "staging wait");
        if (k + STAGE_SLOTS - 1 < nchunk)
            if (const int rc = issue(k + STAGE_SLOTS - 1); rc != F4J_OK) return rc;
        const int64_t r0 = k * rpc, n = std::min(rpc, rows - r0);
        host_rows_copy((char *)g.slot[sl], nullptr, (char *)host + r0 * pitch, n, row_bytes, pitch);
    }
    return F4J_OK;
}

bool valid_op(int op, bool cplx) { return op == F4J_OP_N || op == F4J_OP_T || (cplx && op == F4J_OP_C); }

// Shape/argument validation shared by every GEMM entry.  Sizes in elements.
int validate(bool cplx, int opA, int opB, int64_t M, int64_t N, int64_t K, const void *A, int64_t lda, const void *B,
             int64_t ldb, const void *C, int64_t ldc, int precision) {
    if (!valid_op(opA, cplx) || !valid_op(opB, cplx))
        return fail(F4J_ERR_INVALID_ARG, "invalid op code (N=0, T=1" + std::string(cplx ? ", C=2)" : ")"));
    if (M < 0 || N < 0 || K < 0) return fail(F4J_ERR_INVALID_ARG, "negative matrix dimension");
    if (precision < F4J_PREC_AUTO || precision > F4J_PREC_SPLIT_INT8)
        return fail(F4J_ERR_INVALID_ARG, "invalid precision code " + std::to_string(precision));
    const int64_t rowA = opA == F4J_OP_N ? K : M, rowB = opB == F4J_OP_N ? N : K;
    const int64_t nrA = opA == F4J_OP_N ? M : K, nrB = opB == F4J_OP_N ? K : N;
    if (M > 0 && N > 0) {
        if (!C) return fail(F4J_ERR_INVALID_ARG, "C is null");
        if (ldc < std::max<int64_t>(N, 1)) return fail(F4J_ERR_INVALID_ARG, "ldc < N");
    }
    if (M > 0 && N > 0 && K > 0) {
        if (!A || !B) return fail(F4J_ERR_INVALID_ARG, "A or B is null");
        if (lda < std::max<int64_t>(rowA, 1)) return fail(F4J_ERR_INVALID_ARG, "lda smaller than the stored row length of A");
        if (ldb < std::max<int64_t>(rowB, 1)) return fail(F4J_ERR_INVALID_ARG, "ldb smaller than the stored row length of B");
    }
    // Flag4j indexes with 32-bit ints (M:arrays/Shape.java:397-412): larger
    // matrices cannot come from it, but the library itself is 64-bit clean.
    (void)nrA;
    (void)nrB;
    return F4J_OK;
}

// Tile configuration per shape bucket: the 128x128 tile when it yields enough
// workgroups to fill the 256 CUs, else the 64x64 tile.
constexpr int NUM_TILE_CFGS = NUM_FP64_TILES;

int pick_tile(int64_t M, int64_t Nx) {
    {
        std::lock_guard<std::mutex> g(S().cfg_mu);
        if (S().cfg.tile_config >= 0) return S().cfg.tile_config;
    }
    // 128x128 tiles with 8 waves (4 per SIMD) once they cover the CUs;
    // 64x64 tiles below that.
    const int64_t t128 = ((M + 127) / 128) * ((Nx + 127) / 128);
    return t128 >= 256 ? 4 : 1;
}

std::string tile_name(int cfg) {
    switch (cfg) {
#define F4J_TILE_NAME(id, BM, BN, BK, WGM, WGN, MINW) \
    case id:                                        \
        return "MFMA_F64_" #BM "x" #BN "x" #BK "_w" #WGM "x" #WGN "_o" #MINW;
        F4J_FP64_TILES(F4J_TILE_NAME)
#undef F4J_TILE_NAME
        default: return "MFMA_F64_?";
    }
}

int resolve_precision(int precision, int64_t M, int64_t N, int64_t K, int opA, int opB) {
    if (precision != F4J_PREC_AUTO) return precision;
    // Entries held by A and B (ReDeMatMultDispatcher.java:158: a.dataLength() + b.dataLength()).
    (void)opA;
    (void)opB;
    const int64_t total = M * K + K * N;
    std::lock_guard<std::mutex> g(S().cfg_mu);
    if (total < S().cfg.small_threshold) return F4J_PREC_STRICT;
    return S().cfg.default_precision;
}

OpKind kind_of(int op, bool cplx) {
    if (!cplx) return op == F4J_OP_N ? R_N : R_T;
    return op == F4J_OP_N ? Z_N : (op == F4J_OP_T ? Z_T : Z_C);
}

int default_slices_for(int prec) {
    std::lock_guard<std::mutex> g(S().cfg_mu);
    if (S().cfg.default_slices > 0) return S().cfg.default_slices;
    return prec == F4J_PREC_SPLIT_FP16 ? 5 : 6;
}

// Whether `slices` (the caller's value) leaves the count to the library: then a
// product whose residual check fails is computed once more with a count raised
// by what the measured error asks for, before any FP64 fallback.
bool auto_slices(int slices) {
    if (slices > 0) return false;
    std::lock_guard<std::mutex> g(S().cfg_mu);
    return S().cfg.default_slices == 0;
}

// Digits to add so that an error `err` falls below `tol`: each digit adds
// `bits_per` bits; one bit of margin.
int extra_digits(double err, double tol, int bits_per) {
    const double bits = std::log2(err / tol) + 1.0;
    return std::max(1, (int)std::ceil(bits / bits_per));
}

// Split-precision real GEMM (synchronous: its workspace is shared and its
// residual check reads back one number).
// With `retry` set, a failed check whose error more digits can fix returns F4J_OK
// with *retry = the digit count to run again with (no fallback, no report).
int run_split(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B,
              int64_t ldb, double *C, int64_t ldc, int prec, int slices, hipStream_t stream, int *retry) {
    const int carrier = prec == F4J_PREC_SPLIT_FP16 ? SPLIT_FP16 : SPLIT_BF16;
    const int S_ = slices > 0 ? slices : default_slices_for(prec);
    if (S_ < 1 || S_ > SPLIT_MAX_SLICES)
        return fail(F4J_ERR_INVALID_ARG, "slices must be in 1.." + std::to_string(SPLIT_MAX_SLICES));
    Config c;
    {
        std::lock_guard<std::mutex> g(S().cfg_mu);
        c = S().cfg;
    }
    const SplitPlan pl = plan_split(M, N, K, S_, carrier, c.split_check_rows);
    State &s = S();
    std::lock_guard<std::mutex> g(s.ws_mu);
    int rc = grow(s.ws, pl.bytes);
    if (rc != F4J_OK) return rc;
    hipEvent_t *ev = c.split_timing ? s.tev : nullptr;
    F4J_HIP(launch_split_gemm(pl, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, s.ws.p, stream, ev ? ev[0] : nullptr,
                              ev ? ev[1] : nullptr),
            "split gemm launch");
    f4j_split_report rep{};
    rep.product_ms = -1.0;
    rep.precision = prec;
    rep.slices = S_;
    rep.products = pl.products();
    rep.launches = (int)pl.groups.size();
    rep.checked_rows = (int)std::min<int64_t>(c.split_check_rows, M);
    rep.max_rel_err = -1.0;
    // NaN or Inf among the inputs (flagged by the exponent kernels): the digit
    // slices cannot carry them; FP64 arithmetic propagates them as the reference's does.
    // The check's kernels go first, then both read-backs into pinned memory: one wait.
    if (rep.checked_rows > 0)
        F4J_HIP(launch_split_residual(pl, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, rep.checked_rows, s.ws.p, stream),
                "split residual launch");
    F4J_HIP(hipMemcpyAsync(&s.rb->nonfinite, (char *)s.ws.p + pl.off_flag, sizeof(int), hipMemcpyDeviceToHost, stream),
            "copy flag");
    if (rep.checked_rows > 0)
        F4J_HIP(hipMemcpyAsync(&s.rb->err, (char *)s.ws.p + pl.off_err, sizeof(double), hipMemcpyDeviceToHost, stream),
                "copy residual");
    F4J_HIP(hipStreamSynchronize(stream), "stream synchronize");
    const int nonfinite = s.rb->nonfinite;
    const double err = rep.checked_rows > 0 ? s.rb->err : 0.0;
    if (nonfinite) {
        GemmProblem p{M, N, K, A, lda, B, ldb, C, ldc, 0};
        F4J_HIP(launch_gemm_f64(opA ? R_T : R_N, opB ? R_T : R_N, p, stream, pick_tile(M, N)),
                "fp64 for non-finite inputs");
        rep.precision = F4J_PREC_FP64;
    } else if (rep.checked_rows > 0) {
        rep.max_rel_err = err;
        if (!(err <= c.split_tolerance)) {
            if (retry && S_ < SPLIT_MAX_SLICES && std::isfinite(err) && c.split_tolerance > 0) {
                *retry = std::min(SPLIT_MAX_SLICES,
                                  S_ + extra_digits(err, c.split_tolerance, carrier == SPLIT_FP16 ? 10 : 9));
                return F4J_OK;
            }
            if (!c.split_fallback) {
                t_split = rep;
                return fail(F4J_ERR_ACCURACY, "split product error " + std::to_string(err) + " above tolerance " +
                                                  std::to_string(c.split_tolerance));
            }
            GemmProblem p{M, N, K, A, lda, B, ldb, C, ldc, 0};
            F4J_HIP(launch_gemm_f64(opA ? R_T : R_N, opB ? R_T : R_N, p, stream, pick_tile(M, N)), "fp64 fallback");
            rep.precision = F4J_PREC_FP64;
        }
    }
    F4J_HIP(hipStreamSynchronize(stream), "stream synchronize");
    if (ev) {
        float ms = 0.f;
        F4J_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime");
        rep.product_ms = ms;
    }
    t_split = rep;
    return F4J_OK;
}

// FP64 MFMA GEMM of real or complex (real-expanded) operands.
hipError_t launch_fp64(bool cplx, int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                       const double *B, int64_t ldb, double *C, int64_t ldc, hipStream_t stream) {
    const int w = cplx ? 2 : 1;
    GemmProblem p{M, N * w, K * w, A, lda * w, B, ldb * w, C, ldc * w, 0};
    return launch_gemm_f64(kind_of(opA, cplx), kind_of(opB, cplx), p, stream, pick_tile(p.M, p.N));
}

// Complex split-precision GEMM: the batched split engine on a batch of one
// (three real products in the 3M form when use3m, else four), with the same
// residual check, fallback and report as the real form.
int run_split_cplx(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B,
                   int64_t ldb, double *C, int64_t ldc, int prec, int slices, int use3m, hipStream_t stream,
                   int *retry) {
    const int carrier = prec == F4J_PREC_SPLIT_FP16 ? SPLIT_FP16 : SPLIT_BF16;
    const int S_ = slices > 0 ? slices : default_slices_for(prec);
    if (S_ < 1 || S_ > SPLIT_MAX_SLICES)
        return fail(F4J_ERR_INVALID_ARG, "slices must be in 1.." + std::to_string(SPLIT_MAX_SLICES));
    Config c;
    {
        std::lock_guard<std::mutex> g(S().cfg_mu);
        c = S().cfg;
    }
    State &s = S();
    std::lock_guard<std::mutex> g(s.ws_mu);
    // The plan of the last call with these arguments is idle (every call waits for
    // its stream before it returns, holding ws_mu): reuse it, else replace it.
    const State::CplxSplitKey key{opA, opB, carrier, S_, use3m != 0, c.split_check_rows, M, N, K, lda, ldb, ldc,
                                  A, B, C};
    if (!s.cplx_split || !(s.cplx_split_key == key)) {
        split_batch_destroy(s.cplx_split);
        s.cplx_split = nullptr;
        const std::vector<BatchProblem> pr{BatchProblem{M, N, K, A, lda, B, ldb, C, ldc}};
        std::string err;
        int status = F4J_OK;
        s.cplx_split = split_batch_create(true, opA, opB, carrier, S_, use3m != 0, c.split_check_rows, pr, &err,
                                          &status);
        if (!s.cplx_split) return fail(status, err);
        s.cplx_split_key = key;
    }
    SplitBatch *b = s.cplx_split;
    hipEvent_t *ev = c.split_timing ? s.tev : nullptr;
    hipError_t e = split_batch_enqueue(b, stream, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr);
    f4j_split_report rep{};
    rep.precision = prec;
    rep.slices = S_;
    rep.products = split_batch_real_products(b);
    rep.launches = 1;
    rep.checked_rows = (int)std::min<int64_t>(c.split_check_rows, M);
    rep.max_rel_err = -1.0;
    rep.product_ms = -1.0;
    double perr = -1.0;
    if (e == hipSuccess) e = split_batch_errors(b, &perr, 1, stream);      // synchronises
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
        split_batch_destroy(s.cplx_split);
        s.cplx_split = nullptr;
        return hip_fail(e, "complex split gemm");
    }
    if (ev) {
        float ms = 0.f;
        F4J_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime");
        rep.product_ms = ms;
    }
    if (std::isnan(perr)) {
        // NaN or Inf among the inputs: the digit slices cannot carry them; FP64
        // arithmetic propagates them as the reference's does.
        F4J_HIP(launch_fp64(true, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, stream), "fp64 for non-finite inputs");
        F4J_HIP(hipStreamSynchronize(stream), "stream synchronize");
        rep.precision = F4J_PREC_FP64;
    } else if (rep.checked_rows > 0) {
        rep.max_rel_err = perr;
        if (!(perr <= c.split_tolerance)) {
            if (retry && S_ < SPLIT_MAX_SLICES && std::isfinite(perr) && c.split_tolerance > 0) {
                *retry = std::min(SPLIT_MAX_SLICES,
                                  S_ + extra_digits(perr, c.split_tolerance, carrier == SPLIT_FP16 ? 10 : 9));
                return F4J_OK;
            }
            if (!c.split_fallback) {
                t_split = rep;
                return fail(F4J_ERR_ACCURACY, "complex split product error " + std::to_string(perr) +
                                                  " above tolerance " + std::to_string(c.split_tolerance));
            }
            F4J_HIP(launch_fp64(true, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, stream), "fp64 fallback");
            F4J_HIP(hipStreamSynchronize(stream), "stream synchronize");
            rep.precision = F4J_PREC_FP64;
        }
    }
    t_split = rep;
    return F4J_OK;
}

// Modular (CRT) split-precision real GEMM on the INT8 matrix cores;
// `slices` = number of moduli (0: enough for FP64 accuracy at this K).
// With `fixed` set the workspace is that buffer (no allocation): products whose
// workspace does not fit are rejected, and a retry is only asked for if it fits.
int run_crt(bool cplx, int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
            const double *B, int64_t ldb, double *C, int64_t ldc, int slices, int use3m, hipStream_t stream,
            int *retry, const DevBuf *fixed) {
    Config c;
    {
        std::lock_guard<std::mutex> g(S().cfg_mu);
        c = S().cfg;
    }
    const int n = slices > 0 ? slices : (c.default_slices > 0 ? c.default_slices : crt_default_moduli(K));
    if (n < 2 || n > CRT_MAX_MODULI)
        return fail(F4J_ERR_INVALID_ARG, "split_int8: moduli must be in 2.." + std::to_string(CRT_MAX_MODULI) +
                                             " (0 = chosen from K), got " + std::to_string(n));
    CrtPlan pl;
    if (K >= (int64_t(1) << 17) || !plan_crt(M, N, K, n, c.split_check_rows, &pl))
        return fail(F4J_ERR_INVALID_ARG, "split_int8: " + std::to_string(n) +
                                             " moduli cannot represent products with K = " + std::to_string(K) +
                                             " (need 2..20 moduli and K < 131072)");
    // Real products: one job.  Complex products op(A) = a + ib, op(B) = c + id
    // (b, d carry the conjugation sign): 3M (Gauss) T1 = ac, T2 = bd,
    // T3 = (a + b)(c + d), Re = T1 - T2, Im = T3 - T1 - T2; or 4M.
    struct Job {
        double ca_a, cb_a, ca_b, cb_b, cre, cim;
    };
    const double sa = opA == F4J_OP_C ? -1.0 : 1.0, sb = opB == F4J_OP_C ? -1.0 : 1.0;
    std::vector<Job> jobs;
    if (!cplx) jobs = {{1, 0, 1, 0, 1, 0}};
    else if (use3m) jobs = {{1, 0, 1, 0, 1, -1}, {0, sa, 0, sb, -1, -1}, {1, sa, 1, sb, 0, 1}};
    else jobs = {{1, 0, 1, 0, 1, 0}, {0, sa, 0, sb, -1, 0}, {1, 0, 0, sb, 0, 1}, {0, sa, 1, 0, 0, 1}};
    // complex: every job's residues and product residues at once (one workspace per job)
    const size_t need = cplx ? jobs.size() * crt_job_stride(pl) : pl.bytes;
    if (fixed && need > fixed->cap)
        return fail(F4J_ERR_INVALID_ARG, "split_int8 without allocation: this product needs " +
                                             std::to_string(need) + " bytes of scratch, the limit is " +
                                             std::to_string(fixed->cap) + " (about 1024^3 with 16 moduli)");
    State &s = S();
    std::lock_guard<std::mutex> g(s.ws_mu);
    if (!fixed) {
        const int rc = grow(s.ws, need);
        if (rc != F4J_OK) return rc;
    }
    void *wsp = fixed ? fixed->p : s.ws.p;
    const bool timing = c.split_timing != 0;
    const int check_rows = (int)std::min<int64_t>(c.split_check_rows, M);
    char *w = (char *)wsp;
    double *part = (double *)(w + pl.off_part);
    // Whatever path leaves this function, the side stream is idle first: its
    // work writes the shared workspace, which the next call may grow.
    struct SideIdle {
        hipStream_t side = nullptr;
        ~SideIdle() {
            if (side) (void)hipStreamSynchronize(side);
        }
    } side_idle;
    side_idle.side = s.side;
    {
        // B's exponents and residues on the side stream, beside A's on this one
        const CrtSide side{s.side, s.split_fork_ev, s.split_join_ev};
        if (!cplx) {
            const CrtOperand oa = crt_operand(A, lda, opA == F4J_OP_N, false);
            const CrtOperand ob = crt_operand(B, ldb, opB != F4J_OP_N, false);
            F4J_HIP(launch_crt_gemm(pl, oa, ob, M, N, K, C, ldc, CrtOutput{}, wsp, stream,
                                    timing ? s.tev[0] : nullptr, timing ? s.tev[1] : nullptr, &side),
                    "crt gemm launch");
        } else {
            // all jobs at once: each operand read once per pass for every job's view
            CrtJobs cj;
            cj.n = (int)jobs.size();
            for (size_t jb = 0; jb < jobs.size(); ++jb) {
                cj.a_ca[jb] = jobs[jb].ca_a, cj.a_cb[jb] = jobs[jb].cb_a;
                cj.b_ca[jb] = jobs[jb].ca_b, cj.b_cb[jb] = jobs[jb].cb_b;
                cj.cre[jb] = jobs[jb].cre, cj.cim[jb] = jobs[jb].cim;
            }
            const CrtOperand oa = crt_operand(A, lda, opA == F4J_OP_N, true);
            const CrtOperand ob = crt_operand(B, ldb, opB != F4J_OP_N, true);
            F4J_HIP(launch_crt_gemm_jobs(pl, oa, ob, M, N, K, C, ldc, cj, wsp, stream, timing ? s.tev[0] : nullptr,
                                         timing ? s.tev[1] : nullptr, &side),
                    "crt gemm launch");
        }
        if (check_rows > 0) {
            // The check's compensated sums need only A and B: they follow B's
            // split on the side stream (which has waited for the inputs) and run
            // beside the residue products.
            F4J_HIP(launch_residual_partial(cplx, opA, opB, M, N, K, A, lda, B, ldb, check_rows, part, s.side),
                    "crt residual launch");
            F4J_HIP(hipEventRecord(s.join_ev, s.side), "hipEventRecord");
        }
    }
    f4j_split_report rep{};
    rep.product_ms = -1.0;
    rep.precision = F4J_PREC_SPLIT_INT8;
    rep.slices = n;
    rep.products = n * (int)jobs.size();
    rep.launches = 1;                        // one product launch covers every job's moduli
    rep.checked_rows = check_rows;
    rep.max_rel_err = -1.0;
    // NaN or Inf among the inputs (flagged by the exponent kernels): the
    // residues cannot carry them, so the product is recomputed in FP64, whose
    // arithmetic propagates them as the reference's does.
    // The check's kernels go first, then both read-backs into pinned memory: one wait.
    if (rep.checked_rows > 0) {
        unsigned long long *eb = (unsigned long long *)(w + pl.off_err);
        F4J_HIP(hipStreamWaitEvent(stream, s.join_ev, 0), "hipStreamWaitEvent");
        F4J_HIP(launch_residual_combine(cplx, M, N, K, C, ldc, rep.checked_rows, part, eb, stream),
                "crt residual launch");
    }
    F4J_HIP(hipMemcpyAsync(&s.rb->nonfinite, w + pl.off_flag, sizeof(int), hipMemcpyDeviceToHost, stream), "copy flag");
    if (rep.checked_rows > 0)
        F4J_HIP(hipMemcpyAsync(&s.rb->err, w + pl.off_err, sizeof(double), hipMemcpyDeviceToHost, stream),
                "copy residual");
    F4J_HIP(hipStreamSynchronize(stream), "stream synchronize");
    const int nonfinite = s.rb->nonfinite;
    const double err = rep.checked_rows > 0 ? s.rb->err : 0.0;
    if (nonfinite) {
        F4J_HIP(launch_fp64(cplx, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, stream), "fp64 for non-finite inputs");
        rep.precision = F4J_PREC_FP64;
    } else if (rep.checked_rows > 0) {
        rep.max_rel_err = err;
        if (!(err <= c.split_tolerance)) {
            if (retry && n < CRT_MAX_MODULI && std::isfinite(err) && c.split_tolerance > 0) {
                // t grows by about 4 bits per modulus: ask for the bits the error lacks
                const int bits = (int)std::ceil(std::log2(err / c.split_tolerance)) + 1;
                const int m = std::min(CRT_MAX_MODULI, std::max(n + 1, crt_moduli_for_bits(crt_bits(n, K) + bits, K)));
                CrtPlan p2;
                if (!fixed || (plan_crt(M, N, K, m, c.split_check_rows, &p2) && p2.bytes <= fixed->cap)) {
                    *retry = m;
                    return F4J_OK;
                }
            }
            if (!c.split_fallback) {
                t_split = rep;
                return fail(F4J_ERR_ACCURACY, "split_int8 product error " + std::to_string(err) +
                                                  " above tolerance " + std::to_string(c.split_tolerance));
            }
            F4J_HIP(launch_fp64(cplx, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, stream), "fp64 fallback");
            rep.precision = F4J_PREC_FP64;
        }
    }
    F4J_HIP(hipStreamSynchronize(stream), "stream synchronize");
    if (timing) {
        float ms = 0.f;
        F4J_HIP(hipEventElapsedTime(&ms, s.tev[0], s.tev[1]), "hipEventElapsedTime");
        rep.product_ms = ms;
    }
    t_split = rep;
    return F4J_OK;
}

// Matrix-vector kind the reference's dispatch gives y = A x (A M x K): small
// products (M K + K < smallThreshold, ReDeMatVecMultDispatcher.java:144-151,
// ReDeMatMultDispatcher.java:158-168) and Matrix.multTranspose with one row
// (RealDenseMatMultTranspose.java:53-82) are one k-ascending sum per row; larger
// real products take the 64-block form (ReDeMatVecMultDispatcher.java:167-181,
// ReDeMatMultDispatcher.java:183-198).  Complex products are always one chain.
int gemv_kind(bool cplx, bool b_transposed, int64_t M, int64_t K) {
    if (cplx) return GEMV_CPLX;
    if (b_transposed) return GEMV_CHAIN;
    std::lock_guard<std::mutex> g(S().cfg_mu);
    return M * K + K < S().cfg.small_threshold ? GEMV_CHAIN : GEMV_BLOCKED;
}

// Launch y = A x in the reference's order; false when the shape is outside the
// matrix-vector kernels' limits (the caller then takes another path).
bool try_gemv(int kind, int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx, double *y,
              int64_t incy, bool conjx, hipStream_t stream, hipError_t *err) {
    if (!gemv_applicable(kind, K, A, lda, x, incx)) return false;
    *err = launch_gemv(kind, M, K, A, lda, x, incx, y, incy, conjx, stream);
    return true;
}

// Device-level GEMM, all pointers on the device, async on `stream`.
// `fixed` (optional): the modular form's workspace, instead of the grown one.
int gemm_dev(bool cplx, int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
             const double *B, int64_t ldb, double *C, int64_t ldc, int precision, int slices, int use3m,
             hipStream_t stream, const DevBuf *fixed = nullptr) {
    if (M == 0 || N == 0) return F4J_OK;
    const int w = cplx ? 2 : 1;
    if (K == 0) {
        F4J_HIP(launch_zero_fill(C, M, N * w, ldc * w, stream), "zero fill");
        return F4J_OK;
    }
    const int prec = resolve_precision(precision, M, N, K, opA, opB);
    // N = 1 with A stored M x K: the reference dispatches a matrix-vector kernel
    // (ReDeMatMultDispatcher.java:183-198); so do FP64 and STRICT here, in its order.
    if (N == 1 && opA == F4J_OP_N && (prec == F4J_PREC_STRICT || prec == F4J_PREC_FP64)) {
        hipError_t e = hipSuccess;
        const int kind = gemv_kind(cplx, opB != F4J_OP_N, M, K);
        if (try_gemv(kind, M, K, A, lda, B, opB == F4J_OP_N ? ldb : 1, C, ldc, opB == F4J_OP_C, stream, &e)) {
            F4J_HIP(e, "matrix-vector launch");
            return F4J_OK;
        }
    }
    if (prec == F4J_PREC_STRICT) {
        F4J_HIP(launch_gemm_strict(cplx, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, stream), "strict gemm launch");
        return F4J_OK;
    }
    if (prec == F4J_PREC_SPLIT_BF16 || prec == F4J_PREC_SPLIT_FP16 || prec == F4J_PREC_SPLIT_INT8) {
        // slices = 0: a failed check is run once more with the count its error asks for
        int again = 0;
        int *retry = auto_slices(slices) ? &again : nullptr;
        auto run = [&](int sl, int *rt) {
            if (prec == F4J_PREC_SPLIT_INT8)
                return run_crt(cplx, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, sl, use3m, stream, rt, fixed);
            if (cplx) return run_split_cplx(opA, opB, M, N, K, A, lda, B, ldb, C, ldc, prec, sl, use3m, stream, rt);
            return run_split(opA, opB, M, N, K, A, lda, B, ldb, C, ldc, prec, sl, stream, rt);
        };
        const int rc = run(slices, retry);
        if (rc != F4J_OK || again == 0) return rc;
        return run(again, nullptr);
    }
    if (prec != F4J_PREC_FP64)
        return fail(F4J_ERR_UNSUPPORTED, "precision " + std::to_string(prec) + " is not available for " +
                                             (cplx ? "complex" : "real") + " products in this build");
    // Tile loads address a block with 32-bit buffer offsets (at most 256 stored
    // rows of a tile): leading dimensions beyond ~1M doubles are not supported.
    if ((lda * w) > (int64_t)(0x7FFFFFF0 / (256 * 8)) || (ldb * w) > (int64_t)(0x7FFFFFF0 / (256 * 8)))
        return fail(F4J_ERR_UNSUPPORTED, "leading dimension too large for the MFMA kernels (limit 1048575 doubles)");
    GemmProblem p{};
    p.M = M;
    p.N = N * w;
    p.K = K * w;
    p.A = A;
    p.lda = lda * w;
    p.B = B;
    p.ldb = ldb * w;
    p.C = C;
    p.ldc = ldc * w;
    F4J_HIP(launch_gemm_f64(kind_of(opA, cplx), kind_of(opB, cplx), p, stream, pick_tile(p.M, p.N)), "fp64 gemm launch");
    return F4J_OK;
}

// Host-array GEMM: operands move in through a staging context's pinned ring, the
// product runs on the context's stream, C moves out through the ring.
int gemm_host(bool cplx, int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
              const double *B, int64_t ldb, double *C, int64_t ldc, int precision, int slices, int use3m) {
    int rc = validate(cplx, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, precision);
    if (rc != F4J_OK) return rc;
    if (M == 0 || N == 0) return F4J_OK;
    rc = ensure_ready();
    if (rc != F4J_OK) return rc;
    std::unique_lock<std::mutex> lk;
    Stage &g = stage_acquire(lk);
    if ((rc = stage_setup(g)) != F4J_OK) return rc;
    const size_t es = cplx ? 16 : 8;
    const int64_t rA = opA == F4J_OP_N ? M : K, cA = opA == F4J_OP_N ? K : M;
    const int64_t rB = opB == F4J_OP_N ? K : N, cB = opB == F4J_OP_N ? N : K;
    if (K > 0) {
        if ((rc = grow(g.A, (size_t)(rA * cA) * es)) != F4J_OK) return rc;
        if ((rc = grow(g.B, (size_t)(rB * cB) * es)) != F4J_OK) return rc;
    }
    if ((rc = grow(g.C, (size_t)(M * N) * es)) != F4J_OK) return rc;
    if (K > 0) {
        if ((rc = stage_in(g, g.B.p, B, rB, cB * es, ldb * es)) != F4J_OK) return rc;
        if ((rc = stage_in(g, g.A.p, A, rA, cA * es, lda * es)) != F4J_OK) return rc;
    }
    rc = gemm_dev(cplx, opA, opB, M, N, K, (const double *)g.A.p, cA, (const double *)g.B.p, cB, (double *)g.C.p, N,
                  precision, slices, use3m, g.st);
    if (rc != F4J_OK) {
        (void)hipStreamSynchronize(g.st);
        return rc;
    }
    return stage_out(g, C, ldc * es, g.C.p, M, N * es);
}

const char *prec_names[] = {"fp64", "strict", "split_bf16", "split_fp16", "split_int8"};

int parse_precision(const std::string &v) {
    for (int i = 0; i < 5; ++i)
        if (v == prec_names[i]) return i;
    return -100;
}

std::string trim(const std::string &x) {
    size_t a = x.find_first_not_of(" \t\r\n"), b = x.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
}

int config_set_locked(Config &c, const std::string &key, const std::string &value) {
    char *end = nullptr;
    if (key == "smallThreshold") {
        const long long v = std::strtoll(value.c_str(), &end, 10);
        if (end == value.c_str() || *end || v < 0) return fail(F4J_ERR_INVALID_ARG, "smallThreshold: not a non-negative integer");
        c.small_threshold = v;
    } else if (key == "defaultPrecision") {
        const int p = parse_precision(value);
        if (p < 0) return fail(F4J_ERR_INVALID_ARG, "defaultPrecision: unknown value '" + value + "'");
        c.default_precision = p;
    } else if (key == "tileConfig") {
        const long v = std::strtol(value.c_str(), &end, 10);
        if (end == value.c_str() || *end || v < -1 || v >= NUM_TILE_CFGS)
            return fail(F4J_ERR_INVALID_ARG, "tileConfig: expected -1.." + std::to_string(NUM_TILE_CFGS - 1));
        c.tile_config = (int)v;
    } else if (key == "splitCheckRows") {
        const long v = std::strtol(value.c_str(), &end, 10);
        if (end == value.c_str() || *end || v < 0 || v > 65536) return fail(F4J_ERR_INVALID_ARG, "splitCheckRows: expected 0..65536");
        c.split_check_rows = (int)v;
    } else if (key == "splitTolerance") {
        const double v = std::strtod(value.c_str(), &end);
        if (end == value.c_str() || *end || !(v >= 0)) return fail(F4J_ERR_INVALID_ARG, "splitTolerance: expected a non-negative number");
        c.split_tolerance = v;
    } else if (key == "splitFallback") {
        if (value != "0" && value != "1") return fail(F4J_ERR_INVALID_ARG, "splitFallback: expected 0 or 1");
        c.split_fallback = value == "1";
    } else if (key == "splitTiming") {
        if (value != "0" && value != "1") return fail(F4J_ERR_INVALID_ARG, "splitTiming: expected 0 or 1");
        c.split_timing = value == "1";
    } else if (key == "tunerCache") {
        c.tuner_cache = value;
    } else if (key == "defaultSlices") {
        const long v = std::strtol(value.c_str(), &end, 10);
        if (end == value.c_str() || *end || v < 0 || v > CRT_MAX_MODULI)
            return fail(F4J_ERR_INVALID_ARG, "defaultSlices: expected 0.." + std::to_string(CRT_MAX_MODULI));
        c.default_slices = (int)v;
    } else {
        return fail(F4J_ERR_INVALID_ARG, "unknown configuration key '" + key + "'");
    }
    return F4J_OK;
}

}  // namespace

void set_error(const std::string &msg) { t_err = msg; }

// Mixed product on the device: mode 0 C = A B with A real, B complex; mode 1
// A complex, B real (RealFieldDenseMatMult, M:linalg/ops/dense/real_field_ops/
// RealFieldDenseMatMult.java:58-1090).  Sizes and leading dimensions in elements.
int mixed_dev(int mode, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
              double *C, int64_t ldc, int precision, int slices, hipStream_t stream) {
    if (M == 0 || N == 0) return F4J_OK;
    if (K == 0) {
        F4J_HIP(launch_zero_fill(C, M, 2 * N, 2 * ldc, stream), "zero fill");
        return F4J_OK;
    }
    const int prec = resolve_precision(precision, M, N, K, F4J_OP_N, F4J_OP_N);
    if (prec == F4J_PREC_STRICT) {
        F4J_HIP(launch_gemm_strict_mixed(mode, M, N, K, A, lda, B, ldb, C, ldc, stream), "strict mixed launch");
        return F4J_OK;
    }
    if (N == 1 && prec == F4J_PREC_FP64) {
        // matrix-vector forms: one k-ascending sum per component, as in the reference
        hipError_t e = hipSuccess;
        if (!try_gemv(mode == 0 ? GEMV_RX : GEMV_CX, M, K, A, lda, B, ldb, C, ldc, false, stream, &e))
            e = launch_gemm_strict_mixed(mode, M, N, K, A, lda, B, ldb, C, ldc, stream);
        F4J_HIP(e, "mixed matrix-vector launch");
        return F4J_OK;
    }
    if (mode == 0) {
        // Each component of C takes the same products and sums as a real product with
        // B's interleaved storage viewed as a K x 2N real matrix.
        return gemm_dev(false, F4J_OP_N, F4J_OP_N, M, 2 * N, K, A, lda, B, 2 * ldb, C, 2 * ldc, prec, slices, 0,
                        stream);
    }
    // mode 1: [Re A; Im A] (2M x K) B (K x N) = [Re C; Im C], then interleave.
    State &s = S();
    std::lock_guard<std::mutex> g(s.mix_mu);
    const size_t wbytes = (size_t)(2 * M * K) * 8, vbytes = (size_t)(2 * M * N) * 8;
    int rc = grow(s.mixws, wbytes + vbytes);
    if (rc != F4J_OK) return rc;
    double *W = (double *)s.mixws.p, *V = (double *)((char *)s.mixws.p + wbytes);
    F4J_HIP(launch_split_planes(A, lda, M, K, W, stream), "split planes");
    rc = gemm_dev(false, F4J_OP_N, F4J_OP_N, 2 * M, N, K, W, K, B, ldb, V, N, prec, slices, 0, stream);
    if (rc == F4J_OK) {
        F4J_HIP(launch_merge_planes(V, M, N, C, ldc, stream), "merge planes");
    }
    // the planes buffer is shared: it is free again only once the stream has used it
    const hipError_t e = hipStreamSynchronize(stream);
    if (rc != F4J_OK) return rc;
    F4J_HIP(e, "stream synchronize");
    return F4J_OK;
}

// Host-array mixed product: A, B and C move through a staging context.
int mixed_host(int mode, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
               double *C, int64_t ldc, int precision, int slices) {
    int rc = validate(false, F4J_OP_N, F4J_OP_N, M, N, K, A, lda, B, ldb, C, ldc, precision);
    if (rc != F4J_OK || M == 0 || N == 0) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    std::unique_lock<std::mutex> lk;
    Stage &g = stage_acquire(lk);
    if ((rc = stage_setup(g)) != F4J_OK) return rc;
    const size_t ea = mode == 1 ? 16 : 8, eb = mode == 0 ? 16 : 8;
    if (K > 0) {
        if ((rc = grow(g.A, (size_t)(M * K) * ea)) != F4J_OK) return rc;
        if ((rc = grow(g.B, (size_t)(K * N) * eb)) != F4J_OK) return rc;
    }
    if ((rc = grow(g.C, (size_t)(M * N) * 16)) != F4J_OK) return rc;
    if (K > 0) {
        if ((rc = stage_in(g, g.B.p, B, K, N * eb, ldb * eb)) != F4J_OK) return rc;
        if ((rc = stage_in(g, g.A.p, A, M, K * ea, lda * ea)) != F4J_OK) return rc;
    }
    rc = mixed_dev(mode, M, N, K, (const double *)g.A.p, K, (const double *)g.B.p, N, (double *)g.C.p, N, precision,
                   slices, g.st);
    if (rc != F4J_OK) {
        (void)hipStreamSynchronize(g.st);
        return rc;
    }
    return stage_out(g, C, ldc * 16, g.C.p, M, N * 16);
}

// y = A x on the device (sizes and strides in elements).
int gemv_dev(bool cplx, int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx, double *y,
             int64_t incy, int precision, hipStream_t stream) {
    if (M == 0) return F4J_OK;
    const int w = cplx ? 2 : 1;
    if (K == 0) {
        F4J_HIP(launch_zero_fill(y, M, w, incy * w, stream), "zero fill");
        return F4J_OK;
    }
    const int kind = gemv_kind(cplx, false, M, K);
    hipError_t e = hipSuccess;
    if (try_gemv(kind, M, K, A, lda, x, incx, y, incy, false, stream, &e)) {
        F4J_HIP(e, "matrix-vector launch");
        return F4J_OK;
    }
    if (kind != GEMV_BLOCKED) {
        // a single k-ascending chain per row is also what the strict kernel computes
        F4J_HIP(launch_gemm_strict(cplx, F4J_OP_N, F4J_OP_N, M, 1, K, A, lda, x, incx, y, incy, stream),
                "strict matrix-vector launch");
        return F4J_OK;
    }
    return fail(F4J_ERR_UNSUPPORTED, "matrix-vector product beyond the kernel limits (slab of 8 rows or x over 2 GiB)");
}

int validate_gemv(bool cplx, int64_t M, int64_t K, const void *A, int64_t lda, const void *x, int64_t incx,
                  const void *y, int64_t incy, int precision) {
    (void)cplx;
    if (M < 0 || K < 0) return fail(F4J_ERR_INVALID_ARG, "negative dimension");
    if (precision < F4J_PREC_AUTO || precision > F4J_PREC_SPLIT_INT8)
        return fail(F4J_ERR_INVALID_ARG, "invalid precision code " + std::to_string(precision));
    if (precision != F4J_PREC_AUTO && precision != F4J_PREC_FP64 && precision != F4J_PREC_STRICT)
        return fail(F4J_ERR_UNSUPPORTED, "matrix-vector products run in FP64 (AUTO, FP64 or STRICT): "
                                         "they are bound by reading A once");
    if (incx < 1 || incy < 1) return fail(F4J_ERR_INVALID_ARG, "incx and incy must be >= 1");
    if (M > 0 && !y) return fail(F4J_ERR_INVALID_ARG, "y is null");
    if (M > 0 && K > 0) {
        if (!A || !x) return fail(F4J_ERR_INVALID_ARG, "A or x is null");
        if (lda < K) return fail(F4J_ERR_INVALID_ARG, "lda < K");
    }
    return F4J_OK;
}

// Host-array y = A x: A, x and y move through a staging context.
int gemv_host(bool cplx, int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx, double *y,
              int64_t incy, int precision) {
    int rc = validate_gemv(cplx, M, K, A, lda, x, incx, y, incy, precision);
    if (rc != F4J_OK || M == 0) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    std::unique_lock<std::mutex> lk;
    Stage &g = stage_acquire(lk);
    if ((rc = stage_setup(g)) != F4J_OK) return rc;
    const size_t es = cplx ? 16 : 8;
    if (K > 0) {
        if ((rc = grow(g.A, (size_t)(M * K) * es)) != F4J_OK) return rc;
        if ((rc = grow(g.B, (size_t)K * es)) != F4J_OK) return rc;
    }
    if ((rc = grow(g.C, (size_t)M * es)) != F4J_OK) return rc;
    if (K > 0) {
        if ((rc = stage_in(g, g.B.p, x, K, es, incx * es)) != F4J_OK) return rc;
        if ((rc = stage_in(g, g.A.p, A, M, K * es, lda * es)) != F4J_OK) return rc;
    }
    rc = gemv_dev(cplx, M, K, (const double *)g.A.p, K, (const double *)g.B.p, 1, (double *)g.C.p, 1, precision, g.st);
    if (rc != F4J_OK) {
        (void)hipStreamSynchronize(g.st);
        return rc;
    }
    return stage_out(g, y, incy * es, g.C.p, M, es);
}

int validate_lstsq(int64_t M, int64_t N, int64_t R, const void *A, int64_t lda, const void *B, int64_t ldb,
                   const void *X, int64_t ldx, int precision) {
    if (M < 0 || N < 0 || R < 0) return fail(F4J_ERR_INVALID_ARG, "negative dimension");
    if (precision < F4J_PREC_AUTO || precision > F4J_PREC_SPLIT_INT8)
        return fail(F4J_ERR_INVALID_ARG, "invalid precision code " + std::to_string(precision));
    // R of the reduced QR is min(M, N) x N: the back solver needs it square
    // (ValidateParameters.ensureSquare), and an empty one has no det.
    if (N == 0 || M < N)
        return fail(F4J_ERR_SHAPE, "least squares needs M >= N >= 1 (R of the reduced QR of a " + std::to_string(M) +
                                       "x" + std::to_string(N) + " matrix must be square and non-empty)");
    if (!A) return fail(F4J_ERR_INVALID_ARG, "A is null");
    if (lda < N) return fail(F4J_ERR_INVALID_ARG, "lda < N");
    // the update and transpose grids take 32 rows per workgroup row, at most 65535 of them
    if (M > (int64_t)65535 * 32)
        return fail(F4J_ERR_UNSUPPORTED, "least squares supports at most 2097120 rows");
    if (R > 0) {
        if (!B || !X) return fail(F4J_ERR_INVALID_ARG, "B or X is null");
        if (ldb < R || ldx < R) return fail(F4J_ERR_INVALID_ARG, "ldb or ldx < R");
    }
    return F4J_OK;
}

// Least squares on device operands (RealLstsqSolver.solve): QR of A, Q, Qh = Q^T
// materialised as the reference's Q.H(), Y = Qh B through gemm_dev at
// `precision`, back substitution.  Waits for `stream` to check det R.
int lstsq_dev(int64_t M, int64_t N, int64_t R, const double *A, int64_t lda, const double *B, int64_t ldb, double *X,
              int64_t ldx, int precision, hipStream_t stream) {
    State &s = S();
    std::lock_guard<std::mutex> lk(s.ls_mu);
    // T, Q, Qh, v, w, qf, diag, Y, det, xs; each part starts on a 256-byte boundary
    const int64_t parts[10] = {M * N, M * N, N * M, M, N, N, N, N * R, 1, N * R};
    int64_t off[10], total = 0;
    for (int i = 0; i < 10; ++i) {
        off[i] = total;
        total += (parts[i] + 31) & ~int64_t(31);
    }
    int rc = grow(s.lsws, (size_t)total * sizeof(double));
    if (rc != F4J_OK) return rc;
    double *base = (double *)s.lsws.p;
    double *T = base + off[0], *Q = base + off[1], *Qh = base + off[2], *v = base + off[3], *w = base + off[4],
           *qf = base + off[5], *diag = base + off[6], *Y = base + off[7], *det = base + off[8], *xs = base + off[9];
    F4J_HIP(hipMemcpy2DAsync(T, N * sizeof(double), A, lda * sizeof(double), N * sizeof(double), M,
                             hipMemcpyDeviceToDevice, stream),
            "copy of A");
    F4J_HIP(launch_householder_qr(T, M, N, N, v, w, qf, diag, stream), "Householder QR launch");
    F4J_HIP(launch_householder_q(T, M, N, N, qf, Q, N, v, w, stream), "Q launch");
    F4J_HIP(launch_transpose(Q, M, N, N, Qh, M, stream), "transpose launch");
    if (R > 0) {
        rc = gemm_dev(false, F4J_OP_N, F4J_OP_N, N, R, M, Qh, M, B, ldb, Y, R, precision, 0, 0, stream);
        if (rc != F4J_OK) return rc;
    }
    F4J_HIP(launch_back_solve(T, N, N, Y, R, X, ldx, R, xs, det, stream), "back substitution launch");
    double d = 0.0;
    F4J_HIP(hipMemcpyAsync(&d, det, sizeof(double), hipMemcpyDeviceToHost, stream), "copy of det R");
    F4J_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    const double ad = std::fabs(d);
    // RealBackSolver / BackSolver.checkSingular: |det| <= EPS_F64 max(rows, cols), or NaN
    if (ad <= 0x1.0p-53 * (double)N || std::isnan(ad)) return fail(F4J_ERR_SINGULAR, "Could not solve system.");
    return F4J_OK;
}

// Host-array least squares: A and B move in through a staging context, X out.
int lstsq_host(int64_t M, int64_t N, int64_t R, const double *A, int64_t lda, const double *B, int64_t ldb, double *X,
               int64_t ldx, int precision) {
    int rc = validate_lstsq(M, N, R, A, lda, B, ldb, X, ldx, precision);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    std::unique_lock<std::mutex> lk;
    Stage &g = stage_acquire(lk);
    if ((rc = stage_setup(g)) != F4J_OK) return rc;
    const int64_t Rs = std::max<int64_t>(R, 1);
    if ((rc = grow(g.A, (size_t)(M * N) * sizeof(double))) != F4J_OK) return rc;
    if ((rc = grow(g.B, (size_t)(M * Rs) * sizeof(double))) != F4J_OK) return rc;
    if ((rc = grow(g.C, (size_t)(N * Rs) * sizeof(double))) != F4J_OK) return rc;
    if ((rc = stage_in(g, g.A.p, A, M, N * sizeof(double), lda * sizeof(double))) != F4J_OK) return rc;
    if (R > 0 && (rc = stage_in(g, g.B.p, B, M, R * sizeof(double), ldb * sizeof(double))) != F4J_OK) return rc;
    rc = lstsq_dev(M, N, R, (const double *)g.A.p, N, (const double *)g.B.p, Rs, (double *)g.C.p, Rs, precision, g.st);
    if (rc != F4J_OK && rc != F4J_ERR_SINGULAR) {
        (void)hipStreamSynchronize(g.st);
        return rc;
    }
    if (R > 0) {
        const std::string msg = t_err;
        const int rc_out = stage_out(g, X, ldx * sizeof(double), g.C.p, N, R * sizeof(double));
        if (rc_out != F4J_OK) return rc_out;
        t_err = msg;
    }
    return rc;
}

int validate_pinv(int64_t M, int64_t N, const void *A, int64_t lda, const void *X, int64_t ldx) {
    if (M < 0 || N < 0) return fail(F4J_ERR_INVALID_ARG, "negative dimension");
    if (M == 0 || N == 0) return fail(F4J_ERR_SHAPE, "pseudo-inverse of an empty matrix (S has no entries)");
    if (!A || !X) return fail(F4J_ERR_INVALID_ARG, "A or X is null");
    if (lda < N) return fail(F4J_ERR_INVALID_ARG, "lda < N");
    if (ldx < M) return fail(F4J_ERR_INVALID_ARG, "ldx < M");
    if (std::max(M, N) > (int64_t)65535 * 32) return fail(F4J_ERR_UNSUPPORTED, "pseudo-inverse supports at most 2097120 rows or columns");
    return F4J_OK;
}

// Pseudo-inverse on device operands (Invert.pInv): for a tall A (M >= N)
// X = R^-1 Q^T, the least-squares solve with B = I_M (Q^T I = Q^T exactly, so
// the product is skipped); a wide A goes through its transpose, X = ((A^T)^+)^T.
// Waits for `stream` to check det R.
int pinv_dev(int64_t M, int64_t N, const double *A, int64_t lda, double *X, int64_t ldx, hipStream_t stream) {
    State &s = S();
    std::lock_guard<std::mutex> lk(s.ls_mu);
    const bool tall = M >= N;
    const int64_t m = tall ? M : N, n = tall ? N : M;       // the tall matrix T0 (m x n) that is factored
    // T, Q, Qh, v, w, qf, diag, P (n x m: A^+ of T0, for a wide A), det, xs
    const int64_t parts[10] = {m * n, m * n, n * m, m, n, n, n, tall ? 0 : n * m, 1, n * m};
    int64_t off[10], total = 0;
    for (int i = 0; i < 10; ++i) {
        off[i] = total;
        total += (parts[i] + 31) & ~int64_t(31);
    }
    int rc = grow(s.lsws, (size_t)total * sizeof(double));
    if (rc != F4J_OK) return rc;
    double *base = (double *)s.lsws.p;
    double *T = base + off[0], *Q = base + off[1], *Qh = base + off[2], *v = base + off[3], *w = base + off[4],
           *qf = base + off[5], *diag = base + off[6], *P = base + off[7], *det = base + off[8], *xs = base + off[9];
    if (tall)
        F4J_HIP(hipMemcpy2DAsync(T, n * sizeof(double), A, lda * sizeof(double), n * sizeof(double), m,
                                 hipMemcpyDeviceToDevice, stream),
                "copy of A");
    else
        F4J_HIP(launch_transpose(A, M, N, lda, T, n, stream), "transpose of A");
    F4J_HIP(launch_householder_qr(T, m, n, n, v, w, qf, diag, stream), "Householder QR launch");
    F4J_HIP(launch_householder_q(T, m, n, n, qf, Q, n, v, w, stream), "Q launch");
    F4J_HIP(launch_transpose(Q, m, n, n, Qh, m, stream), "transpose launch");
    // R Y = Qh, m right-hand sides: Y = T0^+ (n x m)
    F4J_HIP(launch_back_solve(T, n, n, Qh, m, tall ? X : P, tall ? ldx : m, m, xs, det, stream),
            "back substitution launch");
    if (!tall) F4J_HIP(launch_transpose(P, n, m, m, X, ldx, stream), "transpose of (A^T)^+");
    double d = 0.0;
    F4J_HIP(hipMemcpyAsync(&d, det, sizeof(double), hipMemcpyDeviceToHost, stream), "copy of det R");
    F4J_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    const double ad = std::fabs(d);
    // Invert.invDiag: |prod s_i| <= EPS_F64 k (|det R| = prod s_i), or NaN
    if (ad <= 0x1.0p-53 * (double)n || std::isnan(ad)) return fail(F4J_ERR_SINGULAR, "Could not invert.");
    return F4J_OK;
}

// Host-array pseudo-inverse: A moves in through a staging context, X out.
int pinv_host(int64_t M, int64_t N, const double *A, int64_t lda, double *X, int64_t ldx) {
    int rc = validate_pinv(M, N, A, lda, X, ldx);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    std::unique_lock<std::mutex> lk;
    Stage &g = stage_acquire(lk);
    if ((rc = stage_setup(g)) != F4J_OK) return rc;
    if ((rc = grow(g.A, (size_t)(M * N) * sizeof(double))) != F4J_OK) return rc;
    if ((rc = grow(g.C, (size_t)(N * M) * sizeof(double))) != F4J_OK) return rc;
    if ((rc = stage_in(g, g.A.p, A, M, N * sizeof(double), lda * sizeof(double))) != F4J_OK) return rc;
    rc = pinv_dev(M, N, (const double *)g.A.p, N, (double *)g.C.p, M, g.st);
    if (rc != F4J_OK && rc != F4J_ERR_SINGULAR) {
        (void)hipStreamSynchronize(g.st);
        return rc;
    }
    const std::string msg = t_err;
    const int rc_out = stage_out(g, X, ldx * sizeof(double), g.C.p, N, M * sizeof(double));
    if (rc_out != F4J_OK) return rc_out;
    t_err = msg;
    return rc;
}

int runtime_ready() { return ensure_ready(); }

hipStream_t runtime_stream(void *stream) { return pick_stream(stream); }

}  // namespace f4j

using namespace f4j;

extern "C" {

int f4j_abi_version(void) { return F4J_ABI_VERSION; }

int f4j_init(int device) {
    State &s = S();
    std::lock_guard<std::mutex> g(s.mu);
    const bool was = s.inited;
    const int rc = init_locked(s, device);
    if (rc == F4J_OK && !was) load_tuner_cache();
    return rc;
}

int f4j_shutdown(void) {
    State &s = S();
    std::lock_guard<std::mutex> g(s.mu);
    if (!s.inited) return F4J_OK;
    F4J_HIP(hipSetDevice(s.device), "hipSetDevice");
    F4J_HIP(hipStreamSynchronize(s.stream), "stream synchronize");
    F4J_HIP(hipStreamSynchronize(s.side), "stream synchronize");
    for (Stage &st : s.stages) {
        std::lock_guard<std::mutex> sg(st.mu);
        stage_release(st);
    }
    {
        std::lock_guard<std::mutex> wg(s.ws_mu);
        split_batch_destroy(s.cplx_split);
        s.cplx_split = nullptr;
    }
    if (s.rb) {
        F4J_HIP(hipHostFree(s.rb), "hipHostFree");
        s.rb = nullptr;
    }
    for (DevBuf *b : {&s.ws, &s.mixws, &s.lsws})
        if (b->p) {
            F4J_HIP(hipFree(b->p), "hipFree");
            b->p = nullptr;
            b->cap = 0;
        }
    for (hipEvent_t &e : s.tev)
        if (e) {
            F4J_HIP(hipEventDestroy(e), "hipEventDestroy");
            e = nullptr;
        }
    for (hipEvent_t *e : {&s.fork_ev, &s.join_ev, &s.split_fork_ev, &s.split_join_ev})
        if (*e) {
            F4J_HIP(hipEventDestroy(*e), "hipEventDestroy");
            *e = nullptr;
        }
    F4J_HIP(hipStreamDestroy(s.side), "hipStreamDestroy");
    s.side = nullptr;
    F4J_HIP(hipStreamDestroy(s.stream), "hipStreamDestroy");
    s.stream = nullptr;
    s.inited = false;
    s.device = -1;
    return F4J_OK;
}

const char *f4j_last_error(void) { return t_err.c_str(); }

int f4j_device_info_get(f4j_device_info *out) {
    if (!out) return fail(F4J_ERR_INVALID_ARG, "out is null");
    int rc = ensure_ready();
    if (rc != F4J_OK) return rc;
    const hipDeviceProp_t &p = S().prop;
    std::memset(out, 0, sizeof(*out));
    out->device = S().device;
    out->compute_units = p.multiProcessorCount;
    out->clock_khz = p.clockRate;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, S().device) != hipSuccess) lds = 0;
    out->lds_bytes_per_cu = lds;
    out->xcds = 8;
    out->hbm_bytes = (int64_t)p.totalGlobalMem;
    // MI355X published dense peaks (FP64 matrix = FP64 vector on CDNA4).
    out->peak_fp64_tflops = 78.6;
    out->peak_bf16_tflops = 2500.0;
    out->peak_hbm_gbs = 8000.0;
    std::snprintf(out->name, sizeof(out->name), "%s", p.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
    return F4J_OK;
}

int f4j_dev_alloc(size_t bytes, void **dptr) {
    if (!dptr) return fail(F4J_ERR_INVALID_ARG, "dptr is null");
    int rc = ensure_ready();
    if (rc != F4J_OK) return rc;
    *dptr = nullptr;
    F4J_HIP(hipMalloc(dptr, bytes ? bytes : 1), "hipMalloc");
    return F4J_OK;
}

int f4j_dev_free(void *dptr) {
    if (!dptr) return F4J_OK;
    int rc = ensure_ready();
    if (rc != F4J_OK) return rc;
    F4J_HIP(hipFree(dptr), "hipFree");
    return F4J_OK;
}

int f4j_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    int rc = ensure_ready();
    if (rc != F4J_OK) return rc;
    if (bytes == 0) return F4J_OK;
    if (!dst || !src) return fail(F4J_ERR_INVALID_ARG, "null pointer");
    F4J_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, pick_stream(stream)), "hipMemcpyAsync H2D");
    return F4J_OK;
}

int f4j_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    int rc = ensure_ready();
    if (rc != F4J_OK) return rc;
    if (bytes == 0) return F4J_OK;
    if (!dst || !src) return fail(F4J_ERR_INVALID_ARG, "null pointer");
    hipStream_t st = pick_stream(stream);
    F4J_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
    F4J_HIP(hipStreamSynchronize(st), "stream synchronize");
    return F4J_OK;
}

int f4j_stream_sync(void *stream) {
    int rc = ensure_ready();
    if (rc != F4J_OK) return rc;
    F4J_HIP(hipStreamSynchronize(pick_stream(stream)), "stream synchronize");
    return F4J_OK;
}

int f4j_dgemm(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B,
              int64_t ldb, double *C, int64_t ldc, int precision, int slices) {
    return gemm_host(false, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, precision, slices, 0);
}

int f4j_dgemm_dev(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B,
                  int64_t ldb, double *C, int64_t ldc, int precision, int slices, void *stream) {
    int rc = validate(false, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, precision);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    return gemm_dev(false, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, precision, slices, 0, pick_stream(stream));
}

int f4j_dgemm_dev_noalloc(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                          const double *B, int64_t ldb, double *C, int64_t ldc, int precision, int slices,
                          void *stream) {
    int rc = validate(false, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, precision);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    State &s = S();
    const int prec = resolve_precision(precision, M, N, K, opA, opB);
    if (prec == F4J_PREC_SPLIT_BF16 || prec == F4J_PREC_SPLIT_FP16)
        return fail(F4J_ERR_UNSUPPORTED, "f4j_dgemm_dev_noalloc: the digit forms need a workspace; use FP64, "
                                         "STRICT or SPLIT_INT8");
    DevBuf fixed;
    fixed.p = s.device < NOALLOC_MAX_DEVICES ? s.scratch_ptr[s.device] : nullptr;
    fixed.cap = NOALLOC_SCRATCH_BYTES;
    if (!fixed.p) return fail(F4J_ERR_UNSUPPORTED, "f4j_dgemm_dev_noalloc: no scratch on device " +
                                                       std::to_string(s.device));
    return gemm_dev(false, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, prec, slices, 0, pick_stream(stream), &fixed);
}

int f4j_zgemm(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B,
              int64_t ldb, double *C, int64_t ldc, int precision, int slices, int use_3m) {
    return gemm_host(true, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, precision, slices, use_3m);
}

int f4j_zgemm_dev(int opA, int opB, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B,
                  int64_t ldb, double *C, int64_t ldc, int precision, int slices, int use_3m, void *stream) {
    int rc = validate(true, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, precision);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    return gemm_dev(true, opA, opB, M, N, K, A, lda, B, ldb, C, ldc, precision, slices, use_3m, pick_stream(stream));
}

int f4j_dzgemm(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
               int64_t ldc, int precision, int slices) {
    return mixed_host(0, M, N, K, A, lda, B, ldb, C, ldc, precision, slices);
}

int f4j_dzgemm_dev(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                   double *C, int64_t ldc, int precision, int slices, void *stream) {
    int rc = validate(false, F4J_OP_N, F4J_OP_N, M, N, K, A, lda, B, ldb, C, ldc, precision);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    return mixed_dev(0, M, N, K, A, lda, B, ldb, C, ldc, precision, slices, pick_stream(stream));
}

int f4j_zdgemm(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
               int64_t ldc, int precision, int slices) {
    return mixed_host(1, M, N, K, A, lda, B, ldb, C, ldc, precision, slices);
}

int f4j_zdgemm_dev(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                   double *C, int64_t ldc, int precision, int slices, void *stream) {
    int rc = validate(false, F4J_OP_N, F4J_OP_N, M, N, K, A, lda, B, ldb, C, ldc, precision);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    return mixed_dev(1, M, N, K, A, lda, B, ldb, C, ldc, precision, slices, pick_stream(stream));
}

int f4j_dgemv(int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx, double *y,
              int64_t incy, int precision) {
    return gemv_host(false, M, K, A, lda, x, incx, y, incy, precision);
}

int f4j_dgemv_dev(int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx, double *y,
                  int64_t incy, int precision, void *stream) {
    int rc = validate_gemv(false, M, K, A, lda, x, incx, y, incy, precision);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    return gemv_dev(false, M, K, A, lda, x, incx, y, incy, precision, pick_stream(stream));
}

int f4j_dlstsq(int64_t M, int64_t N, int64_t R, const double *A, int64_t lda, const double *B, int64_t ldb, double *X,
               int64_t ldx, int precision) {
    return lstsq_host(M, N, R, A, lda, B, ldb, X, ldx, precision);
}

int f4j_dlstsq_dev(int64_t M, int64_t N, int64_t R, const double *A, int64_t lda, const double *B, int64_t ldb,
                   double *X, int64_t ldx, int precision, void *stream) {
    int rc = validate_lstsq(M, N, R, A, lda, B, ldb, X, ldx, precision);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    return lstsq_dev(M, N, R, A, lda, B, ldb, X, ldx, precision, pick_stream(stream));
}

int f4j_dpinv(int64_t M, int64_t N, const double *A, int64_t lda, double *X, int64_t ldx) {
    return pinv_host(M, N, A, lda, X, ldx);
}

int f4j_dpinv_dev(int64_t M, int64_t N, const double *A, int64_t lda, double *X, int64_t ldx, void *stream) {
    int rc = validate_pinv(M, N, A, lda, X, ldx);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    return pinv_dev(M, N, A, lda, X, ldx, pick_stream(stream));
}

int f4j_zgemv(int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx, double *y,
              int64_t incy, int precision) {
    return gemv_host(true, M, K, A, lda, x, incx, y, incy, precision);
}

int f4j_zgemv_dev(int64_t M, int64_t K, const double *A, int64_t lda, const double *x, int64_t incx, double *y,
                  int64_t incy, int precision, void *stream) {
    int rc = validate_gemv(true, M, K, A, lda, x, incx, y, incy, precision);
    if (rc != F4J_OK) return rc;
    if ((rc = ensure_ready()) != F4J_OK) return rc;
    return gemv_dev(true, M, K, A, lda, x, incx, y, incy, precision, pick_stream(stream));
}

int f4j_last_split_report(f4j_split_report *out) {
    if (!out) return fail(F4J_ERR_INVALID_ARG, "out is null");
    *out = t_split;
    return F4J_OK;
}

int f4j_config_set(const char *key, const char *value) {
    if (!key || !value) return fail(F4J_ERR_INVALID_ARG, "null key or value");
    State &s = S();
    int rc;
    {
        std::lock_guard<std::mutex> g(s.cfg_mu);
        rc = config_set_locked(s.cfg, trim(key), trim(value));
    }
    if (rc == F4J_OK && trim(key) == "tunerCache") load_tuner_cache();   // a cache named after init is read now
    return rc;
}

int f4j_config_get(const char *key, char *buf, size_t buflen) {
    if (!key || !buf || buflen == 0) return fail(F4J_ERR_INVALID_ARG, "null key or buffer");
    State &s = S();
    std::string v;
    {
        std::lock_guard<std::mutex> g(s.cfg_mu);
        const std::string k = key;
        if (k == "smallThreshold") v = std::to_string(s.cfg.small_threshold);
        else if (k == "defaultPrecision") v = prec_names[s.cfg.default_precision];
        else if (k == "defaultSlices") v = std::to_string(s.cfg.default_slices);
        else if (k == "tileConfig") v = std::to_string(s.cfg.tile_config);
        else if (k == "splitCheckRows") v = std::to_string(s.cfg.split_check_rows);
        else if (k == "splitTolerance") {
            char b[32];
            std::snprintf(b, sizeof(b), "%.17g", s.cfg.split_tolerance);
            v = b;
        } else if (k == "splitFallback") v = std::to_string(s.cfg.split_fallback);
        else if (k == "splitTiming") v = std::to_string(s.cfg.split_timing);
        else if (k == "tunerCache") v = s.cfg.tuner_cache;
        else return fail(F4J_ERR_INVALID_ARG, "unknown configuration key '" + k + "'");
    }
    std::snprintf(buf, buflen, "%s", v.c_str());
    return F4J_OK;
}

int f4j_config_load(const char *path) {
    if (!path) return fail(F4J_ERR_INVALID_ARG, "null path");
    std::ifstream in(path);
    if (!in) return fail(F4J_ERR_INVALID_ARG, std::string("cannot open ") + path);
    State &s = S();
    std::unique_lock<std::mutex> g(s.cfg_mu);
    Config c = s.cfg;       // all-or-nothing: apply only if every line parses
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#' || t[0] == '!') continue;
        const size_t eq = t.find_first_of("=:");
        if (eq == std::string::npos)
            return fail(F4J_ERR_INVALID_ARG, std::string(path) + ":" + std::to_string(lineno) + ": expected key=value");
        const int rc = config_set_locked(c, trim(t.substr(0, eq)), trim(t.substr(eq + 1)));
        if (rc != F4J_OK) {
            t_err = std::string(path) + ":" + std::to_string(lineno) + ": " + t_err;
            return rc;
        }
    }
    const bool new_cache = c.tuner_cache != s.cfg.tuner_cache;
    s.cfg = c;
    g.unlock();
    if (new_cache) load_tuner_cache();
    return F4J_OK;
}

int f4j_dispatch_name(int is_complex, int opA, int opB, int64_t M, int64_t N, int64_t K, int precision, char *buf,
                      size_t buflen) {
    if (!buf || buflen == 0) return fail(F4J_ERR_INVALID_ARG, "null buffer");
    const bool cplx = is_complex != 0;
    if (!valid_op(opA, cplx) || !valid_op(opB, cplx) || M < 0 || N < 0 || K < 0)
        return fail(F4J_ERR_INVALID_ARG, "invalid op or size");
    if (precision < F4J_PREC_AUTO || precision > F4J_PREC_SPLIT_INT8)
        return fail(F4J_ERR_INVALID_ARG, "invalid precision code");
    std::string name;
    const int prec = resolve_precision(precision, M, N, K, opA, opB);
    if (M == 0 || N == 0) name = "EMPTY";
    else if (K == 0) name = "ZERO_FILL";
    else if (N == 1 && opA == F4J_OP_N && (prec == F4J_PREC_STRICT || prec == F4J_PREC_FP64)) {
        const int kind = gemv_kind(cplx, opB != F4J_OP_N, M, K);
        name = kind == GEMV_BLOCKED ? "GEMV_BLOCKED" : (kind == GEMV_CHAIN ? "GEMV_CHAIN" : "GEMV_COMPLEX");
    } else if (prec == F4J_PREC_STRICT) name = "STRICT";
    else if (prec == F4J_PREC_FP64) name = tile_name(pick_tile(M, N * (cplx ? 2 : 1)));
    else name = std::string("SPLIT_") + prec_names[prec];
    std::snprintf(buf, buflen, "%s", name.c_str());
    return F4J_OK;
}

}  // extern "C"
