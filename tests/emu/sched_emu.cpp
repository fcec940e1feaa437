// Host check of the XCD-aware tile schedule (flag4j_amd/csrc/xcd_schedule.h)
// used by the grouped FP64 and batched split plans.
//
// For random batches of products (tiles with a product id and a cost) it
// checks: every tile is dealt exactly once; the tiles of a product whose cost
// is at most a quarter of an XCD's fair share all sit on workgroups of one XCD
// (w % 8, or w itself with fewer than 8 workgroups); and the busiest
// workgroup's load is within the list-scheduling bound of the two-level deal
//     max load <= total / workers + max tile cost + max unit cost / floor(workers / 8)
// (Graham's bound per level, XCDs weighted by their workgroup counts).  A planted mistake
// (tiles of one product spread round-robin over all workgroups) must fail the
// XCD check.
//
// usage: sched_emu <products> <workers> <seed>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../flag4j_amd/csrc/xcd_schedule.h"

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

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const int nprod = std::atoi(argv[1]), workers = std::atoi(argv[2]);
    std::mt19937 g((unsigned)std::atoi(argv[3]));
    std::uniform_int_distribution<int> tiles_of(1, 64), kdim(1, 40);
    std::vector<double> cost;
    std::vector<int> group;
    std::vector<double> pcost(nprod, 0.0);
    for (int p = 0; p < nprod; ++p) {
        const int nt = (p == 0 && nprod > 1) ? 4000 : tiles_of(g);     // product 0: a large one
        const double c = 16.0 * (kdim(g) + 4);
        for (int t = 0; t < nt; ++t) {
            cost.push_back(c);
            group.push_back(p);
            pcost[p] += c;
        }
    }
    double total = 0.0, cmax = 0.0;
    for (double c : cost) {
        total += c;
        cmax = std::fmax(cmax, c);
    }
    const int nxcd = workers < 8 ? workers : 8;
    const double share = total / (4.0 * nxcd);
    double umax = cmax;
    for (double c : pcost)
        if (c <= share) umax = std::fmax(umax, c);

    auto check = [&](const std::vector<std::vector<int>> &per, bool expect_ok) {
        const int before = g_fail;
        std::vector<int> seen(cost.size(), 0), xcd_of(nprod, -1);
        double maxload = 0.0;
        bool spread = false;
        for (int w = 0; w < (int)per.size(); ++w) {
            double load = 0.0;
            for (int i : per[w]) {
                CHECK(i >= 0 && i < (int)cost.size(), "tile index %d out of range", i);
                ++seen[i];
                load += cost[i];
                const int p = group[i];
                if (pcost[p] <= share) {
                    if (xcd_of[p] < 0) xcd_of[p] = w % nxcd;
                    else if (xcd_of[p] != w % nxcd) spread = true;
                }
            }
            maxload = std::fmax(maxload, load);
        }
        for (size_t i = 0; i < cost.size(); ++i) CHECK(seen[i] == 1, "tile %zu dealt %d times", i, seen[i]);
        if (expect_ok) {
            CHECK(!spread, "a small product's tiles sit on more than one XCD");
            const double bound = total / workers + cmax + umax / (double)(workers / nxcd);
            CHECK(maxload <= bound * (1 + 1e-12), "max load %.1f above the bound %.1f", maxload, bound);
        }
        return spread || g_fail > before;
    };
    check(f4j::xcd_schedule(cost, group, workers), true);
    // planted mistake: round-robin over all workgroups spreads products over XCDs
    if (workers >= 16 && nprod > 1) {
        std::vector<std::vector<int>> rr(workers);
        for (size_t i = 0; i < cost.size(); ++i) rr[i % workers].push_back((int)i);
        const int saved = g_fail;
        const bool caught = check(rr, false);
        g_fail = saved;
        CHECK(caught, "the XCD check misses a round-robin deal");
    }
    if (g_fail) {
        std::printf("FAILED %d checks\n", g_fail);
        return 1;
    }
    std::printf("OK products=%d workers=%d tiles=%zu\n", nprod, workers, cost.size());
    return 0;
}
