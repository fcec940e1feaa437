// Grouped (batched) GEMM: thousands of independent products of uneven shapes
// in persistent launches captured once into a hipGraph.
//
// No reference counterpart: Flag4j multiplies one pair per call (each call
// pays a dispatcher lookup and a thread-pool fork/join,
// M:linalg/ops/dispatch/ReDeMatMultDispatcher.java:157-171,
// M:concurrency/ThreadManager.java:155-183; M: = src/main/java/org/flag4j/).
// This is the batched entry point that replaces a Java loop of Matrix.mult /
// A.H().mult(B) calls.
//
// Plan.  Problems are split into tile-shape buckets: each problem goes to the
// shape (64 x 64, 128 x 64, 64 x 128 or 128 x 128) with the least padded work
// per unit of measured efficiency, so a 400 x 40 product takes 128 x 64 tiles
// rather than 128 x 128.  Each bucket runs one persistent launch of the FP64
// MFMA tile kernel with the bucket's tile configuration; its tiles are assigned to workgroups by
// longest-processing-time-first list scheduling on an estimated cost, so the
// workgroups finish together, with each product's tiles kept on one XCD (one
// L2 fetch of its operands).  The launches are captured into one hipGraph at
// plan creation; running the plan replays the graph.  The autotuner times each
// candidate tile configuration per bucket on the plan's own problems and
// records the winner per (device, operand kinds, bucket) in a cache that can
// be saved to and loaded from a text file.
//
// Split-precision plans (f4j_grouped_plan_create_split) capture the batched
// split engine of split_batch.hip instead: exponents, digit slices, one
// persistent slice-product launch over every output tile, residual check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <numeric>
#include <queue>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/flag4j_amd.h"
#include "crt_gemm.h"
#include "f4j_internal.h"
#include "gemm_tile.h"
#include "split_batch.h"
#include "split_gemm.h"
#include "xcd_schedule.h"

namespace f4j {
namespace {

constexpr int NUM_BUCKETS = 4;
// Tile shape of each bucket and the fraction of the FP64 MFMA rate its default
// configuration reaches on padded work in batches of 16..512-sized products
// (measured on MI355X; the 64-wide tiles carry more load and barrier overhead
// per MFMA).
constexpr int BUCKET_BM[NUM_BUCKETS] = {64, 128, 64, 128};
constexpr int BUCKET_BN[NUM_BUCKETS] = {64, 64, 128, 128};
constexpr double BUCKET_EFF[NUM_BUCKETS] = {0.50, 0.60, 0.60, 0.68};
// Tile configurations the autotuner tries (ids of F4J_FP64_TILES).
constexpr int CANDIDATES[] = {1, 7, 8, 0, 4};
// Untuned defaults per bucket: the bucket's own tile shape.
constexpr int DEFAULT_CFG[NUM_BUCKETS] = {1, 8, 7, 4};

struct TileGeom {
    int BM, BN, BK, threads;
};

TileGeom geom_of(int cfg) {
    if (cfg >= F4J_DIRECT_BASE) {              // direct-load kernels: one wave per tile, k steps of 4
        static const int FM[F4J_NUM_DIRECT] = {2, 4, 2, 4}, FN[F4J_NUM_DIRECT] = {2, 2, 4, 4};
        const int i = cfg - F4J_DIRECT_BASE;
        return TileGeom{16 * FM[i], 16 * FN[i], 4, 64};
    }
    switch (cfg) {
#define F4J_TILE_GEOM(id, BM, BN, BK, WGM, WGN, MINW) \
    case id: return TileGeom{BM, BN, BK, WGM * WGN * 64};
        F4J_FP64_TILES(F4J_TILE_GEOM)
#undef F4J_TILE_GEOM
        default: return TileGeom{64, 64, 16, 256};
    }
}

// The bucket whose tile shape gives the least padded M x N~ area per unit of
// efficiency (K is padded alike by every shape).  Ties go to the smaller tile.
int bucket_of(int64_t M, int64_t Nx) {
    int best = 0;
    double best_cost = 0.0;
    for (int b = 0; b < NUM_BUCKETS; ++b) {
        const double pm = (double)((M + BUCKET_BM[b] - 1) / BUCKET_BM[b] * BUCKET_BM[b]);
        const double pn = (double)((Nx + BUCKET_BN[b] - 1) / BUCKET_BN[b] * BUCKET_BN[b]);
        const double cost = pm * pn / BUCKET_EFF[b];
        if (b == 0 || cost < best_cost) {
            best = b;
            best_cost = cost;
        }
    }
    return best;
}

// Tuner cache: key "arch kinds bucket" -> cfg.
std::mutex g_tune_mu;
std::map<std::string, int> g_tune;

std::string tune_key(const std::string &arch, int ka, int kb, int bucket) {
    return arch + " " + std::to_string(ka) + "," + std::to_string(kb) + " " + std::to_string(bucket);
}

std::string device_arch() {
    int dev = 0;
    hipDeviceProp_t p{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return "unknown";
    std::string a = p.gcnArchName;
    return a.substr(0, a.find(':'));
}

}  // namespace

// Occupancy of the grouped kernel for a configuration (workgroups per CU).
int grouped_wgs_per_cu(OpKind ka, OpKind kb, int cfg);

}  // namespace f4j

using namespace f4j;

struct f4j_grouped_plan {
    int is_complex = 0, opA = 0, opB = 0, precision = 0;
    OpKind ka = R_N, kb = R_N;
    int64_t count = 0;
    int cus = 0;
    std::vector<GemmProblem> probs;          // real-expanded problems
    std::vector<int> prob_bucket;
    struct Bucket {
        int cfg = -1, wgs = 0;
        int64_t items = 0;
        GroupedItem *d_items = nullptr;
        int *d_begin = nullptr;
    };
    Bucket buckets[NUM_BUCKETS];
    GemmProblem *d_probs = nullptr;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double flops = 0;
    f4j::SplitBatch *split = nullptr;        // split-precision plan (precision SPLIT_*), else FP64 buckets
    int64_t split_tiles = 0;                 // output tiles of the split plan's product launch
    // Split plans with splitFallback = 1: the FP64 buckets recompute, in the same
    // graph, the products whose checked error is above splitTolerance (or whose
    // operands hold NaN / Inf); `guard` selects them on the device.
    int *d_slot = nullptr;
    f4j::GroupedGuard guard{};
    hipStream_t last = nullptr;              // stream of the last run (the split report waits on it)
};

namespace f4j {
namespace {

int fail(int code, const std::string &m) {
    set_error(m);
    return code;
}

int hipfail(hipError_t e, const char *what) {
    return fail((e == hipErrorOutOfMemory) ? F4J_ERR_OUT_OF_MEMORY : F4J_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define G_HIP(call, what)                              \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return hipfail(e_, what); \
    } while (0)

void free_bucket(f4j_grouped_plan::Bucket &b) {
    if (b.d_items) (void)hipFree(b.d_items);
    if (b.d_begin) (void)hipFree(b.d_begin);
    b.d_items = nullptr;
    b.d_begin = nullptr;
    b.items = 0;
    b.wgs = 0;
}

// The direct-load kernels read real operands through 32-bit buffer offsets:
// every problem's A and B must span fewer than 2^31 bytes.
bool direct_ok(const f4j_grouped_plan &pl) {
    if (pl.is_complex) return false;
    for (const GemmProblem &p : pl.probs) {
        const bool a_rows = pl.ka == R_N, b_rows = pl.kb == R_N;
        const int64_t ea = a_rows ? (p.M - 1) * p.lda + p.K : (p.K - 1) * p.lda + p.M;
        const int64_t eb = b_rows ? (p.K - 1) * p.ldb + p.N : (p.N - 1) * p.ldb + p.K;
        if (ea * 8 >= (int64_t(1) << 31) || eb * 8 >= (int64_t(1) << 31)) return false;
    }
    return true;
}

// Build bucket b's LPT schedule for tile configuration cfg and upload it.
int schedule_bucket(f4j_grouped_plan &pl, int b, int cfg) {
    f4j_grouped_plan::Bucket &bk = pl.buckets[b];
    free_bucket(bk);
    bk.cfg = cfg;
    const TileGeom g = geom_of(cfg);
    struct Cand {
        double cost;
        GroupedItem it;
    };
    std::vector<Cand> items;
    for (int64_t i = 0; i < pl.count; ++i) {
        if (pl.prob_bucket[i] != b) continue;
        const GemmProblem &p = pl.probs[i];
        const int64_t tm = (p.M + g.BM - 1) / g.BM, tn = (p.N + g.BN - 1) / g.BN;
        // MFMA time of a full tile (ragged tiles run full MFMAs) plus a fixed
        // prologue/epilogue cost worth about four k-slices.
        const double cost = (double)g.BM * g.BN * (double)(((p.K + g.BK - 1) / g.BK) + 4) * g.BK;
        for (int64_t x = 0; x < tm; ++x)
            for (int64_t y = 0; y < tn; ++y) items.push_back({cost, GroupedItem{(int)i, (int)x, (int)y}});
    }
    bk.items = (int64_t)items.size();
    if (items.empty()) return F4J_OK;
    const int per_cu = grouped_wgs_per_cu(pl.ka, pl.kb, cfg);
    const int wgs = (int)std::min<int64_t>((int64_t)items.size(), (int64_t)pl.cus * std::max(per_cu, 1));
    // tiles dealt per XCD (each product's tiles to one L2), longest first (xcd_schedule.h)
    std::vector<double> cost(items.size());
    std::vector<int> group(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
        cost[i] = items[i].cost;
        group[i] = items[i].it.prob;
    }
    const std::vector<std::vector<int>> idx = xcd_schedule(cost, group, wgs);
    std::vector<std::vector<GroupedItem>> per(wgs);
    for (int w = 0; w < wgs; ++w)
        for (int i : idx[w]) per[w].push_back(items[i].it);
    std::vector<GroupedItem> flat;
    std::vector<int> begin(wgs + 1, 0);
    flat.reserve(items.size());
    for (int w = 0; w < wgs; ++w) {
        begin[w] = (int)flat.size();
        flat.insert(flat.end(), per[w].begin(), per[w].end());
    }
    begin[wgs] = (int)flat.size();
    G_HIP(hipMalloc(&bk.d_items, flat.size() * sizeof(GroupedItem)), "hipMalloc (grouped items)");
    G_HIP(hipMalloc(&bk.d_begin, begin.size() * sizeof(int)), "hipMalloc (grouped schedule)");
    G_HIP(hipMemcpy(bk.d_items, flat.data(), flat.size() * sizeof(GroupedItem), hipMemcpyHostToDevice), "copy items");
    G_HIP(hipMemcpy(bk.d_begin, begin.data(), begin.size() * sizeof(int), hipMemcpyHostToDevice), "copy schedule");
    bk.wgs = wgs;
    return F4J_OK;
}

hipError_t enqueue_bucket(const f4j_grouped_plan &pl, int b, hipStream_t st) {
    const f4j_grouped_plan::Bucket &bk = pl.buckets[b];
    if (bk.items == 0) return hipSuccess;
    return launch_gemm_f64_grouped(pl.ka, pl.kb, pl.d_probs, bk.d_items, bk.d_begin, bk.wgs, bk.cfg, st, pl.guard);
}

int capture(f4j_grouped_plan &pl) {
    if (pl.exec) {
        (void)hipGraphExecDestroy(pl.exec);
        pl.exec = nullptr;
    }
    if (pl.graph) {
        (void)hipGraphDestroy(pl.graph);
        pl.graph = nullptr;
    }
    G_HIP(hipStreamBeginCapture(pl.cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipError_t e = hipSuccess;
    if (pl.split) e = split_batch_enqueue(pl.split, pl.cap);
    if (!pl.split || pl.guard.err)
        for (int b = 0; b < NUM_BUCKETS && e == hipSuccess; ++b) e = enqueue_bucket(pl, b, pl.cap);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(pl.cap, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hipfail(e, "grouped launch during capture");
    }
    if (e2 != hipSuccess) return hipfail(e2, "hipStreamEndCapture");
    pl.graph = g;
    G_HIP(hipGraphInstantiate(&pl.exec, pl.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    return F4J_OK;
}

void destroy_plan(f4j_grouped_plan *pl) {
    if (!pl) return;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    for (auto &b : pl->buckets) free_bucket(b);
    if (pl->d_probs) (void)hipFree(pl->d_probs);
    if (pl->d_slot) (void)hipFree(pl->d_slot);
    split_batch_destroy(pl->split);
    if (pl->cap) (void)hipStreamDestroy(pl->cap);
    delete pl;
}

OpKind kind_for(int op, bool cplx) {
    if (!cplx) return op == F4J_OP_N ? R_N : R_T;
    return op == F4J_OP_N ? Z_N : (op == F4J_OP_T ? Z_T : Z_C);
}

// Argument checks shared by the plan constructors.
int check_problems(bool cplx, int opA, int opB, int64_t count, const int64_t *M, const int64_t *N, const int64_t *K,
                   const double *const *A, const int64_t *lda, const double *const *B, const int64_t *ldb,
                   double *const *C, const int64_t *ldc) {
    auto okop = [&](int op) { return op == F4J_OP_N || op == F4J_OP_T || (cplx && op == F4J_OP_C); };
    if (!okop(opA) || !okop(opB)) return fail(F4J_ERR_INVALID_ARG, "invalid op code");
    if (count < 0) return fail(F4J_ERR_INVALID_ARG, "negative count");
    if (count > 0 && (!M || !N || !K || !A || !lda || !B || !ldb || !C || !ldc))
        return fail(F4J_ERR_INVALID_ARG, "null problem array");
    if (count > (int64_t)0x7FFFFFFF) return fail(F4J_ERR_INVALID_ARG, "count too large");
    const int w = cplx ? 2 : 1;
    for (int64_t i = 0; i < count; ++i) {
        if (M[i] < 0 || N[i] < 0 || K[i] < 0)
            return fail(F4J_ERR_INVALID_ARG, "problem " + std::to_string(i) + ": negative dimension");
        const int64_t rowA = opA == F4J_OP_N ? K[i] : M[i], rowB = opB == F4J_OP_N ? N[i] : K[i];
        const bool empty = M[i] == 0 || N[i] == 0;
        if (!empty && (ldc[i] < std::max<int64_t>(N[i], 1) || !C[i]))
            return fail(F4J_ERR_INVALID_ARG, "problem " + std::to_string(i) + ": bad C or ldc");
        if (!empty && K[i] > 0 &&
            (!A[i] || !B[i] || lda[i] < std::max<int64_t>(rowA, 1) || ldb[i] < std::max<int64_t>(rowB, 1)))
            return fail(F4J_ERR_INVALID_ARG, "problem " + std::to_string(i) + ": bad A/B or leading dimension");
        if (lda[i] * w > (int64_t)(0x7FFFFFF0 / (256 * 8)) || ldb[i] * w > (int64_t)(0x7FFFFFF0 / (256 * 8)))
            return fail(F4J_ERR_UNSUPPORTED, "problem " + std::to_string(i) + ": leading dimension too large");
    }
    return F4J_OK;
}

// The FP64 side of a plan: the real-expanded problem table on the device and
// one scheduled bucket per tile shape (tile configuration from the tuner cache,
// else the bucket's default).  pl.is_complex, opA, opB and count are set.
int setup_fp64(f4j_grouped_plan &pl, const int64_t *M, const int64_t *N, const int64_t *K, const double *const *A,
               const int64_t *lda, const double *const *B, const int64_t *ldb, double *const *C, const int64_t *ldc) {
    const bool cplx = pl.is_complex != 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(F4J_ERR_NO_DEVICE, "no HIP device");
    pl.ka = kind_for(pl.opA, cplx);
    pl.kb = kind_for(pl.opB, cplx);
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return fail(F4J_ERR_HIP, "hipGetDeviceProperties failed");
    pl.cus = prop.multiProcessorCount;
    const int w = cplx ? 2 : 1;
    const int64_t count = pl.count;
    pl.probs.resize(count);
    pl.prob_bucket.resize(count);
    for (int64_t i = 0; i < count; ++i) {
        const bool empty = M[i] == 0 || N[i] == 0;
        GemmProblem &p = pl.probs[i];
        p.M = empty ? 0 : M[i];
        p.N = empty ? 0 : N[i] * w;
        p.K = K[i] * w;
        p.A = A[i];
        p.lda = lda[i] * w;
        p.B = B[i];
        p.ldb = ldb[i] * w;
        p.C = C[i];
        p.ldc = ldc[i] * w;
        p.flags = (vec2_ok(p.A, p.lda, stored_cols(pl.ka, true, p.M, p.N, p.K)) &&
                   vec2_ok(p.B, p.ldb, stored_cols(pl.kb, false, p.M, p.N, p.K)))
                      ? 1
                      : 0;
        pl.prob_bucket[i] = bucket_of(p.M, p.N);
    }
    if (count > 0) {
        G_HIP(hipMalloc(&pl.d_probs, count * sizeof(GemmProblem)), "hipMalloc (grouped problems)");
        G_HIP(hipMemcpy(pl.d_probs, pl.probs.data(), count * sizeof(GemmProblem), hipMemcpyHostToDevice),
              "copy problems");
    }
    const std::string arch = device_arch();
    for (int b = 0; b < NUM_BUCKETS; ++b) {
        int cfg = DEFAULT_CFG[b];
        {
            std::lock_guard<std::mutex> g(g_tune_mu);
            auto it = g_tune.find(tune_key(arch, pl.ka, pl.kb, b));
            if (it != g_tune.end()) cfg = it->second;
        }
        if (cfg >= F4J_DIRECT_BASE && !direct_ok(pl)) cfg = DEFAULT_CFG[b];   // real operands below 2^31 bytes only
        if (const int rc = schedule_bucket(pl, b, cfg); rc != F4J_OK) return rc;
    }
    return F4J_OK;
}

}  // namespace
}  // namespace f4j

extern "C" {

int f4j_grouped_plan_create(int is_complex, int opA, int opB, int64_t count, const int64_t *M, const int64_t *N,
                            const int64_t *K, const double *const *A, const int64_t *lda, const double *const *B,
                            const int64_t *ldb, double *const *C, const int64_t *ldc, int precision,
                            f4j_grouped_plan **out) {
    if (!out) return fail(F4J_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    const bool cplx = is_complex != 0;
    if (precision == F4J_PREC_SPLIT_BF16 || precision == F4J_PREC_SPLIT_FP16 || precision == F4J_PREC_SPLIT_INT8)
        return f4j_grouped_plan_create_split(is_complex, opA, opB, count, M, N, K, A, lda, B, ldb, C, ldc, precision,
                                             0, 1, out);
    if (precision != F4J_PREC_FP64 && precision != F4J_PREC_AUTO)
        return fail(F4J_ERR_UNSUPPORTED, "grouped products: precision must be FP64, SPLIT_BF16, SPLIT_FP16 or SPLIT_INT8");
    if (const int rc = check_problems(cplx, opA, opB, count, M, N, K, A, lda, B, ldb, C, ldc); rc != F4J_OK)
        return rc;
    if (const int rc = runtime_ready(); rc != F4J_OK) return rc;
    f4j_grouped_plan *pl = new f4j_grouped_plan;
    pl->is_complex = is_complex;
    pl->opA = opA;
    pl->opB = opB;
    pl->precision = F4J_PREC_FP64;
    pl->count = count;
    hipError_t e = hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking);
    if (e != hipSuccess) {
        destroy_plan(pl);
        return hipfail(e, "grouped plan setup");
    }
    if (const int rc = setup_fp64(*pl, M, N, K, A, lda, B, ldb, C, ldc); rc != F4J_OK) {
        destroy_plan(pl);
        return rc;
    }
    for (const GemmProblem &p : pl->probs) pl->flops += 2.0 * (double)p.M * (double)p.N * (double)p.K;
    const int rc = capture(*pl);
    if (rc != F4J_OK) {
        destroy_plan(pl);
        return rc;
    }
    *out = pl;
    return F4J_OK;
}

int f4j_grouped_plan_run(f4j_grouped_plan *plan, void *stream) {
    if (!plan) return fail(F4J_ERR_INVALID_ARG, "plan is null");
    if (const int rc = runtime_ready(); rc != F4J_OK) return rc;
    if (!plan->exec) return F4J_OK;
    plan->last = runtime_stream(stream);
    G_HIP(hipGraphLaunch(plan->exec, plan->last), "hipGraphLaunch");
    return F4J_OK;
}

int f4j_grouped_plan_destroy(f4j_grouped_plan *plan) {
    if (plan && plan->cap) (void)hipStreamSynchronize(plan->cap);
    destroy_plan(plan);
    return F4J_OK;
}

int f4j_grouped_plan_info(const f4j_grouped_plan *plan, double *flops, int *tile_cfg, int64_t *items, int nb) {
    if (!plan) return fail(F4J_ERR_INVALID_ARG, "plan is null");
    if (flops) *flops = plan->flops;
    if (plan->split) {                     // one fixed 64 x 64 slice tile; all tiles in "bucket" 0
        for (int b = 0; b < std::min(nb, NUM_BUCKETS); ++b) {
            if (tile_cfg) tile_cfg[b] = -1;
            if (items) items[b] = 0;
        }
        if (items && nb > 0) items[0] = plan->split_tiles;
        return F4J_OK;
    }
    for (int b = 0; b < std::min(nb, NUM_BUCKETS); ++b) {
        if (tile_cfg) tile_cfg[b] = plan->buckets[b].cfg;
        if (items) items[b] = plan->buckets[b].items;
    }
    return F4J_OK;
}

int f4j_grouped_plan_tune(f4j_grouped_plan *plan, int reps) {
    if (!plan) return fail(F4J_ERR_INVALID_ARG, "plan is null");
    if (plan->split) return F4J_OK;        // split plans use the one slice tile; nothing to choose
    if (reps < 1) reps = 3;
    const std::string arch = device_arch();
    hipEvent_t e0, e1;
    G_HIP(hipEventCreate(&e0), "hipEventCreate");
    G_HIP(hipEventCreate(&e1), "hipEventCreate");
    for (int b = 0; b < NUM_BUCKETS; ++b) {
        if (plan->buckets[b].items == 0) continue;
        float best = 1e30f;
        int best_cfg = plan->buckets[b].cfg;
        std::vector<int> cands(std::begin(CANDIDATES), std::end(CANDIDATES));
        if (direct_ok(*plan))                  // the direct-load kernels serve real operands below 2^31 bytes
            for (int i = 0; i < F4J_NUM_DIRECT; ++i) cands.push_back(F4J_DIRECT_BASE + i);
        for (int cfg : cands) {
            int rc = schedule_bucket(*plan, b, cfg);
            if (rc != F4J_OK) return rc;
            G_HIP(enqueue_bucket(*plan, b, plan->cap), "tune warmup");
            G_HIP(hipEventRecord(e0, plan->cap), "hipEventRecord");
            for (int r = 0; r < reps; ++r) G_HIP(enqueue_bucket(*plan, b, plan->cap), "tune launch");
            G_HIP(hipEventRecord(e1, plan->cap), "hipEventRecord");
            G_HIP(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0;
            G_HIP(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            if (ms < best) {
                best = ms;
                best_cfg = cfg;
            }
        }
        int rc = schedule_bucket(*plan, b, best_cfg);
        if (rc != F4J_OK) return rc;
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tune[tune_key(arch, plan->ka, plan->kb, b)] = best_cfg;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // persist the cache where the configuration names one (read back at the next init)
    char path[4096];
    if (f4j_config_get("tunerCache", path, sizeof(path)) == F4J_OK && path[0])
        if (const int rc = f4j_tuner_save(path); rc != F4J_OK) return rc;
    return capture(*plan);
}

int f4j_grouped_plan_create_split(int is_complex, int opA, int opB, int64_t count, const int64_t *M,
                                  const int64_t *N, const int64_t *K, const double *const *A, const int64_t *lda,
                                  const double *const *B, const int64_t *ldb, double *const *C, const int64_t *ldc,
                                  int precision, int slices, int use3m, f4j_grouped_plan **out) {
    if (!out) return fail(F4J_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    const bool cplx = is_complex != 0;
    if (precision != F4J_PREC_SPLIT_BF16 && precision != F4J_PREC_SPLIT_FP16 && precision != F4J_PREC_SPLIT_INT8)
        return fail(F4J_ERR_INVALID_ARG, "split plans take precision SPLIT_BF16, SPLIT_FP16 or SPLIT_INT8");
    const bool int8 = precision == F4J_PREC_SPLIT_INT8;
    if (!int8 && (slices < 0 || slices > SPLIT_MAX_SLICES)) return fail(F4J_ERR_INVALID_ARG, "slices must be in 0..8");
    if (int8 && (slices < 0 || slices > CRT_MAX_MODULI))
        return fail(F4J_ERR_INVALID_ARG, "moduli must be 0 or in 2.." + std::to_string(CRT_MAX_MODULI));
    if (const int rc = check_problems(cplx, opA, opB, count, M, N, K, A, lda, B, ldb, C, ldc); rc != F4J_OK)
        return rc;
    if (const int rc = runtime_ready(); rc != F4J_OK) return rc;
    char buf[32];
    int check_rows = 4, fallback = 1;
    double tolerance = 1e-15;
    if (f4j_config_get("splitCheckRows", buf, sizeof(buf)) == F4J_OK) check_rows = std::atoi(buf);
    if (f4j_config_get("splitFallback", buf, sizeof(buf)) == F4J_OK) fallback = std::atoi(buf);
    if (f4j_config_get("splitTolerance", buf, sizeof(buf)) == F4J_OK) tolerance = std::strtod(buf, nullptr);
    const int carrier = int8 ? SPLIT_INT8 : precision == F4J_PREC_SPLIT_FP16 ? SPLIT_FP16 : SPLIT_BF16;
    // digit forms, 0 = real 7 (BF16) / 6 (FP16) digits, complex 8 / 7: one digit
    // beyond what uniform data needs; short products (K down to 16) of entries
    // spread over 10^+-4 need it in every entry, not only in the checked rows
    // (measured on 2,000 A^H B products, 3M: FP16 6 digits left 15 products above
    // 1e-15, 4 of them with passing checked rows; 7 left none).  Modular form: 0 =
    // chosen from the largest K (split_batch.h).
    const int S = (slices > 0 || int8) ? slices
                  : carrier == SPLIT_FP16 ? (cplx ? 7 : 6)
                                          : (cplx ? 8 : 7);
    std::vector<BatchProblem> probs(count);
    f4j_grouped_plan *pl = new f4j_grouped_plan;
    pl->is_complex = is_complex;
    pl->opA = opA;
    pl->opB = opB;
    pl->precision = precision;
    pl->count = count;
    for (int64_t i = 0; i < count; ++i) {
        probs[i] = BatchProblem{M[i], N[i], K[i], A[i], lda[i], B[i], ldb[i], C[i], ldc[i]};
        pl->flops += (cplx ? 8.0 : 2.0) * (double)M[i] * (double)N[i] * (double)K[i];
    }
    std::string err;
    int status = F4J_OK;
    pl->split = split_batch_create(cplx, opA, opB, carrier, S, use3m != 0, check_rows, probs, &err, &status);
    if (!pl->split) {
        destroy_plan(pl);
        return fail(status, err);
    }
    for (int64_t i = 0; i < count; ++i)
        if (M[i] > 0 && N[i] > 0 && K[i] > 0) pl->split_tiles += ((M[i] + 63) / 64) * ((N[i] + 63) / 64);
    hipError_t e = hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking);
    if (e != hipSuccess) {
        destroy_plan(pl);
        return hipfail(e, "grouped plan setup");
    }
    if (fallback && check_rows > 0 && count > 0) {
        // FP64 recomputation of the products the check rejects, in the same graph
        if (const int rc = setup_fp64(*pl, M, N, K, A, lda, B, ldb, C, ldc); rc != F4J_OK) {
            destroy_plan(pl);
            return rc;
        }
        const std::vector<int> &slots = split_batch_slots(pl->split);
        e = hipMalloc(&pl->d_slot, slots.size() * sizeof(int));
        if (e == hipSuccess)
            e = hipMemcpy(pl->d_slot, slots.data(), slots.size() * sizeof(int), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            destroy_plan(pl);
            return hipfail(e, "grouped plan setup");
        }
        pl->guard.err = split_batch_err_bits(pl->split);
        pl->guard.slot = pl->d_slot;
        pl->guard.above = __builtin_bit_cast(unsigned long long, tolerance);
    }
    const int rc = capture(*pl);
    if (rc != F4J_OK) {
        destroy_plan(pl);
        return rc;
    }
    *out = pl;
    return F4J_OK;
}

int f4j_grouped_plan_split_report(f4j_grouped_plan *plan, double *err, int64_t count, int *products) {
    if (!plan) return fail(F4J_ERR_INVALID_ARG, "plan is null");
    if (!plan->split) return fail(F4J_ERR_INVALID_ARG, "not a split-precision plan");
    if (count > 0 && !err) return fail(F4J_ERR_INVALID_ARG, "err is null");
    if (products) *products = split_batch_real_products(plan->split);
    if (count > 0)
        G_HIP(split_batch_errors(plan->split, err, std::min<int64_t>(count, plan->count),
                                 plan->last ? plan->last : plan->cap),
              "split report");
    return F4J_OK;
}

int f4j_tuner_save(const char *path) {
    if (!path) return fail(F4J_ERR_INVALID_ARG, "null path");
    std::ofstream o(path);
    if (!o) return fail(F4J_ERR_INVALID_ARG, std::string("cannot write ") + path);
    o << "# flag4j_amd grouped-GEMM tuning cache: <arch> <kindA>,<kindB> <bucket> <tile config>\n";
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (const auto &kv : g_tune) o << kv.first << " " << kv.second << "\n";
    return o.good() ? F4J_OK : fail(F4J_ERR_INVALID_ARG, std::string("write failed: ") + path);
}

int f4j_tuner_load(const char *path) {
    if (!path) return fail(F4J_ERR_INVALID_ARG, "null path");
    std::ifstream in(path);
    if (!in) return fail(F4J_ERR_INVALID_ARG, std::string("cannot open ") + path);
    std::map<std::string, int> add;
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string arch, kinds;
        int bucket = -1, cfg = -1;
        if (!(ss >> arch >> kinds >> bucket >> cfg) || bucket < 0 || bucket >= NUM_BUCKETS || cfg < 0 ||
            (cfg >= NUM_FP64_TILES && (cfg < F4J_DIRECT_BASE || cfg >= F4J_DIRECT_BASE + F4J_NUM_DIRECT)))
            return fail(F4J_ERR_INVALID_ARG, std::string(path) + ":" + std::to_string(lineno) + ": malformed entry");
        add[arch + " " + kinds + " " + std::to_string(bucket)] = cfg;
    }
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (const auto &kv : add) g_tune[kv.first] = kv.second;
    return F4J_OK;
}

int f4j_tuner_clear(void) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune.clear();
    return F4J_OK;
}

}  // extern "C"
