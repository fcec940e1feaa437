// Host-side schedule of output tiles onto the workgroups of a persistent
// launch (grouped FP64 plans and batched split plans).
//
// Workgroup w of a launch runs on XCD w % 8: the hardware deals workgroup ids
// round-robin over the 8 XCDs, each with its own L2.  Tiles of one product
// (same `group`) are dealt to a single XCD, so that product's operands are
// fetched into one L2 instead of several; a product above a quarter of an
// XCD's fair share is spread tile by tile instead, so one large product cannot
// unbalance the XCDs.  Units (products, or tiles of the large ones) go longest
// first to the XCD with the least load per workgroup, then each XCD's tiles
// longest first to its least-loaded workgroup (list scheduling, Graham's LPT
// rule).  Every workgroup's load is then at most
//     total / workers + max tile cost + max unit cost / (workgroups of the smallest XCD).
#pragma once
#include <algorithm>
#include <cstdint>
#include <functional>
#include <map>
#include <queue>
#include <utility>
#include <vector>

namespace f4j {

// cost[i], group[i] of tile i; returns the tiles of each of `workers` workgroups.
inline std::vector<std::vector<int>> xcd_schedule(const std::vector<double> &cost, const std::vector<int> &group,
                                                  int workers) {
    using Unit = std::pair<double, std::vector<int>>;
    using Load = std::pair<double, int>;           // (load / capacity, bin)
    // units longest first, each to the bin with the least load per unit of capacity
    auto lpt = [](std::vector<Unit> &units, const std::vector<int> &cap) {
        std::stable_sort(units.begin(), units.end(), [](const Unit &x, const Unit &y) { return x.first > y.first; });
        std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
        for (int k = 0; k < (int)cap.size(); ++k) heap.push({0.0, k});
        std::vector<std::vector<int>> out(cap.size());
        std::vector<double> load(cap.size(), 0.0);
        for (const Unit &u : units) {
            const int k = heap.top().second;
            heap.pop();
            out[k].insert(out[k].end(), u.second.begin(), u.second.end());
            load[k] += u.first;
            heap.push({load[k] / cap[k], k});
        }
        return out;
    };
    std::vector<std::vector<int>> per(std::max(workers, 0));
    if (workers <= 0 || cost.empty()) return per;
    const int nxcd = std::min(8, workers);
    std::map<int, std::vector<int>> of_group;
    double total = 0.0;
    for (size_t i = 0; i < cost.size(); ++i) {
        of_group[group[i]].push_back((int)i);
        total += cost[i];
    }
    const double share = total / (4.0 * nxcd);
    std::vector<Unit> units;
    for (auto &kv : of_group) {
        double c = 0.0;
        for (int i : kv.second) c += cost[i];
        if (c <= share)
            units.push_back({c, kv.second});
        else
            for (int i : kv.second) units.push_back({cost[i], {i}});
    }
    std::vector<int> cap(nxcd);                    // workgroups per XCD (they differ by at most one)
    for (int x = 0; x < nxcd; ++x) cap[x] = (workers - x + nxcd - 1) / nxcd;
    const std::vector<std::vector<int>> xcd = lpt(units, cap);
    for (int x = 0; x < nxcd; ++x) {
        std::vector<int> wk;                       // workgroups on XCD x
        for (int w = x; w < workers; w += nxcd) wk.push_back(w);
        std::vector<Unit> tiles;
        for (int i : xcd[x]) tiles.push_back({cost[i], {i}});
        const std::vector<std::vector<int>> pw = lpt(tiles, std::vector<int>(wk.size(), 1));
        for (size_t k = 0; k < wk.size(); ++k) per[wk[k]] = pw[k];
    }
    return per;
}

}  // namespace f4j
