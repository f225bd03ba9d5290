#include "levelize.h"

#include <algorithm>
#include <numeric>
#include <stdexcept>

namespace grgl_amd {

namespace {

// Groups nodes by level; within a level, longest gather list first, ties by node ID.
void orderByLevel(DirPlan& p, const std::vector<uint32_t>& level, const std::vector<uint64_t>& fanIn,
                  uint32_t numLevels, uint32_t numSamplesFirst) {
    const uint32_t V = p.numNodes;
    p.numLevels = numLevels;
    p.levelStart.assign(numLevels + 1, 0);
    for (uint32_t v = 0; v < V; v++) p.levelStart[level[v] + 1]++;
    for (uint32_t l = 0; l < numLevels; l++) p.levelStart[l + 1] += p.levelStart[l];
    p.order.resize(V);
    {
        std::vector<uint32_t> fill(p.levelStart.begin(), p.levelStart.end() - 1);
        for (uint32_t v = 0; v < V; v++) p.order[fill[level[v]]++] = v;
    }
    for (uint32_t l = 0; l < numLevels; l++) {
        auto b = p.order.begin() + p.levelStart[l], e = p.order.begin() + p.levelStart[l + 1];
        if (l == 0 && numSamplesFirst > 0) {
            // UP level 0: samples first in sample order (they are seeded by a transpose),
            // then the childless non-sample nodes.
            continue;
        }
        std::stable_sort(b, e, [&](uint32_t a, uint32_t c) { return fanIn[a] > fanIn[c]; });
    }
    p.rank.resize(V);
    for (uint32_t j = 0; j < V; j++) p.rank[p.order[j]] = j;
}

// List offsets are kept modulo 2^32 (the traversal reads them relative to a launch's
// first entry), so the total may exceed 2^32; one list may not exceed 2^31 entries.
void checkListLength(uint64_t n) {
    if (n >= (1ULL << 31)) throw std::runtime_error("a gather list exceeds 2^31 entries");
}

std::vector<uint32_t> sortedMutationOrder(const HostGrg& g) {
    // getNodesAndMutations() order: (node, mutId) ascending with unmapped mutations first
    // (grg.cpp:265-301, NodeMutMissLt grg.h:855-865).
    const uint32_t M = g.numMutations(), V = g.numNodes();
    std::vector<uint64_t> start((size_t)V + 2, 0);
    for (uint32_t m = 0; m < M; m++) {
        const uint32_t n = g.mutNode[m];
        start[n == kInvalidNode ? 0 : (size_t)n + 1]++;
    }
    uint64_t acc = 0;
    for (auto& s : start) {
        const uint64_t c = s;
        s = acc;
        acc += c;
    }
    std::vector<uint32_t> order(M);
    for (uint32_t m = 0; m < M; m++) {
        const uint32_t n = g.mutNode[m];
        order[start[n == kInvalidNode ? 0 : (size_t)n + 1]++] = m;
    }
    return order;
}

void fillCoal(DirPlan& p, const HostGrg& g) {
    const uint32_t V = p.numNodes;
    p.coalByRow.resize(V);
    for (uint32_t j = 0; j < V; j++) {
        const uint32_t n = p.order[j];
        p.coalByRow[j] = n < g.numSamples ? 0U : (g.hasCoals() ? g.coal[n] : kCoalNotSet);
    }
}

} // namespace

DirPlan buildUpPlan(const HostGrg& g) {
    validateHostGrg(g);
    DirPlan p;
    const uint32_t V = g.numNodes(), N = g.numSamples;
    p.numNodes = V;
    p.firstPulledLevel = 1;
    std::vector<uint32_t> h(V, 0);
    std::vector<uint64_t> fanIn(V, 0);
    uint32_t maxH = 0;
    for (uint32_t v = N; v < V; v++) {
        const uint64_t b = g.downOff[v], e = g.downOff[v + 1];
        fanIn[v] = e - b;
        if (e > b) {
            uint32_t hv = 0;
            for (uint64_t i = b; i < e; i++) hv = std::max(hv, h[g.downChild[i]]);
            h[v] = hv + 1;
            maxH = std::max(maxH, h[v]);
        }
    }
    orderByLevel(p, h, fanIn, maxH + 1, N);
    p.off.assign(V + 1, 0);
    uint64_t total = 0;
    for (uint32_t j = 0; j < V; j++) {
        const uint32_t n = p.order[j];
        checkListLength(fanIn[n]);
        total += fanIn[n];
        p.off[j + 1] = (uint32_t)total; // modulo 2^32
        p.maxFanIn = std::max<uint32_t>(p.maxFanIn, (uint32_t)fanIn[n]);
    }
    p.src.resize(total);
    uint64_t o = 0;
    for (uint32_t j = 0; j < V; j++) {
        const uint32_t n = p.order[j];
        if (n < N) continue; // samples have no list (fanIn 0)
        for (uint64_t i = g.downOff[n]; i < g.downOff[n + 1]; i++) p.src[o++] = p.rank[g.downChild[i]];
    }
    const uint32_t M = g.numMutations();
    p.outRow.resize(M);
    for (uint32_t m = 0; m < M; m++) p.outRow[m] = g.mutNode[m] == kInvalidNode ? kInvalidNode : p.rank[g.mutNode[m]];
    if (!g.mutMiss.empty()) {
        p.missRow.resize(M);
        for (uint32_t m = 0; m < M; m++) p.missRow[m] = g.mutMiss[m] == kInvalidNode ? kInvalidNode : p.rank[g.mutMiss[m]];
    }
    fillCoal(p, g);
    return p;
}

DirPlan buildDownPlan(const HostGrg& g) {
    validateHostGrg(g);
    DirPlan p;
    const uint32_t V = g.numNodes(), N = g.numSamples, M = g.numMutations();
    p.numNodes = V;
    p.firstPulledLevel = 0;
    // Depth from the roots: a node's row needs every parent's row first.
    std::vector<uint32_t> d(V, 0);
    uint32_t maxD = 0;
    for (uint32_t v = V; v-- > 0;) {
        for (uint64_t i = g.downOff[v]; i < g.downOff[v + 1]; i++) {
            const uint32_t c = g.downChild[i];
            d[c] = std::max(d[c], d[v] + 1);
        }
        maxD = std::max(maxD, d[v]);
    }
    // Gather lists in node space: seeds (tuple order) then parents (descending ID).
    std::vector<uint64_t> cnt(V, 0);
    const bool hasMiss = !g.mutMiss.empty();
    for (uint32_t m = 0; m < M; m++) {
        if (g.mutNode[m] != kInvalidNode) cnt[g.mutNode[m]]++;
        if (hasMiss && g.mutMiss[m] != kInvalidNode) cnt[g.mutMiss[m]]++;
    }
    for (uint64_t i = 0; i < g.numEdges(); i++) cnt[g.downChild[i]]++;
    // Seed rows follow the V node rows: V + m (and V + M + m for missingness), u32 indices.
    if ((uint64_t)V + (uint64_t)M * (hasMiss ? 2 : 1) >= (1ULL << 32))
        throw std::runtime_error("nodes plus mutation seed rows exceed 2^32");
    std::vector<uint64_t> nodeOff(V + 1, 0);
    for (uint32_t v = 0; v < V; v++) {
        checkListLength(cnt[v]);
        nodeOff[v + 1] = nodeOff[v] + cnt[v];
    }
    std::vector<uint32_t> nodeSrc(nodeOff[V]);
    std::vector<uint64_t> cur(nodeOff.begin(), nodeOff.end() - 1);
    const std::vector<uint32_t> tuples = sortedMutationOrder(g);
    for (uint32_t m : tuples) {
        if (g.mutNode[m] != kInvalidNode) nodeSrc[cur[g.mutNode[m]]++] = V + m;
        if (hasMiss && g.mutMiss[m] != kInvalidNode) nodeSrc[cur[g.mutMiss[m]]++] = V + M + m;
    }
    const uint32_t kParentTag = 0x80000000U; // parents are node IDs (< 2^31) until mapped to rows
    for (uint32_t v = V; v-- > 0;) {
        for (uint64_t i = g.downOff[v]; i < g.downOff[v + 1]; i++) nodeSrc[cur[g.downChild[i]]++] = v | kParentTag;
    }
    orderByLevel(p, d, cnt, maxD + 1, 0);
    p.off.assign(V + 1, 0);
    p.src.resize(nodeOff[V]);
    uint64_t o = 0;
    for (uint32_t j = 0; j < V; j++) {
        const uint32_t n = p.order[j];
        for (uint64_t i = nodeOff[n]; i < nodeOff[n + 1]; i++) {
            const uint32_t s = nodeSrc[i];
            p.src[o++] = (s & kParentTag) ? p.rank[s & ~kParentTag] : s;
        }
        p.off[j + 1] = (uint32_t)o; // modulo 2^32
        p.maxFanIn = std::max<uint32_t>(p.maxFanIn, (uint32_t)cnt[n]);
    }
    (void)N;
    p.outRow.resize(g.numSamples);
    for (uint32_t s = 0; s < g.numSamples; s++) p.outRow[s] = p.rank[s];
    fillCoal(p, g);
    return p;
}

uint32_t heavyPrefix(const DirPlan& p, uint32_t lvl, uint32_t heavy) {
    uint32_t lo = p.levelStart[lvl], hi = p.levelStart[lvl + 1];
    const uint32_t begin = lo;
    // Rows in a level are sorted by list length, descending: find the first row <= heavy.
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (p.off[mid + 1] - p.off[mid] > heavy) lo = mid + 1;
        else hi = mid;
    }
    return lo - begin;
}

} // namespace grgl_amd
