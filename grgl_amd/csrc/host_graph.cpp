// Validation and shape statistics of host-side GRGs.
#include "host_graph.h"

#include <algorithm>
#include <stdexcept>
#include <thread>

namespace grgl_amd {

void validateHostGrg(const HostGrg& g) {
    const uint32_t V = g.numNodes();
    if (g.numSamples == 0) throw std::runtime_error("GRG must have at least one sample");
    if (g.ploidy == 0 || g.numSamples % g.ploidy != 0) throw std::runtime_error("numSamples must be a multiple of ploidy");
    if (V < g.numSamples) throw std::runtime_error("numNodes < numSamples");
    if (V >= kInvalidNode) throw std::runtime_error("too many nodes");
    if (g.downOff[0] != 0) throw std::runtime_error("down offsets must start at 0");
    for (uint32_t v = 0; v < V; v++) {
        if (g.downOff[v + 1] < g.downOff[v]) throw std::runtime_error("down offsets must be non-decreasing");
        for (uint64_t e = g.downOff[v]; e < g.downOff[v + 1]; e++) {
            const uint32_t c = g.downChild[e];
            if (c >= v) throw std::runtime_error("nodes are not ordered: child ID >= parent ID");
            if (e > g.downOff[v] && c < g.downChild[e - 1]) throw std::runtime_error("children are not sorted");
        }
    }
    if (g.downChild.size() != g.downOff[V]) throw std::runtime_error("edge array size mismatch");
    for (size_t m = 0; m < g.mutNode.size(); m++) {
        if (g.mutNode[m] != kInvalidNode && g.mutNode[m] >= V) throw std::runtime_error("mutation node out of range");
    }
    if (!g.mutMiss.empty()) {
        if (g.mutMiss.size() != g.mutNode.size()) throw std::runtime_error("missingness map size mismatch");
        for (uint32_t x : g.mutMiss)
            if (x != kInvalidNode && x >= V) throw std::runtime_error("missingness node out of range");
    }
    if (!g.coal.empty() && g.coal.size() != V) throw std::runtime_error("coal count size mismatch");
}

GraphStats computeGraphStats(const HostGrg& g) {
    GraphStats st;
    const uint32_t V = g.numNodes(), N = g.numSamples;
    st.numEdges = g.numEdges();
    std::vector<uint32_t> h(V, 0);
    uint32_t maxH = 0;
    st.fanInHistogram.assign(33, 0);
    for (uint32_t v = N; v < V; v++) {
        uint32_t hv = 0;
        const uint64_t b = g.downOff[v], e = g.downOff[v + 1];
        for (uint64_t i = b; i < e; i++) hv = std::max(hv, h[g.downChild[i]] + 1);
        h[v] = hv;
        maxH = std::max(maxH, hv);
        const uint64_t f = e - b;
        st.maxFanIn = std::max<uint32_t>(st.maxFanIn, (uint32_t)f);
        uint32_t bucket = 0;
        while ((2ULL << bucket) <= f) bucket++;
        st.fanInHistogram[bucket]++;
    }
    st.upLevels = maxH + 1;
    while (st.fanInHistogram.size() > 1 && st.fanInHistogram.back() == 0) st.fanInHistogram.pop_back();
    return st;
}

CarrierParts mutationCarriers(const HostGrg& g, uint32_t threads) {
    validateHostGrg(g);
    const uint32_t M = g.numMutations(), V = g.numNodes(), N = g.numSamples;
    if (threads == 0) threads = std::max(1U, std::thread::hardware_concurrency());
    threads = std::max(1U, std::min<uint32_t>(threads, (M + 255) / 256));
    CarrierParts out;
    out.parts.resize(threads);
    std::vector<std::vector<uint64_t>> counts(threads);
    auto work = [&](uint32_t t) {
        const uint32_t b = (uint32_t)((uint64_t)M * t / threads), e = (uint32_t)((uint64_t)M * (t + 1) / threads);
        std::vector<uint32_t>& part = out.parts[t];
        std::vector<uint32_t> stamp(V, 0), stack;
        std::vector<uint64_t> bits(((size_t)N + 63) / 64, 0); // sorts long lists by a bit scan
        uint32_t epoch = 0;
        for (uint32_t m = b; m < e; m++) {
            const size_t before = part.size();
            const uint32_t root = g.mutNode[m];
            if (root != kInvalidNode) {
                epoch++;
                stack.assign(1, root);
                stamp[root] = epoch;
                while (!stack.empty()) {
                    const uint32_t n = stack.back();
                    stack.pop_back();
                    if (n < N) {
                        part.push_back(n);
                        continue;
                    }
                    for (uint64_t i = g.downOff[n]; i < g.downOff[n + 1]; i++) {
                        const uint32_t c = g.downChild[i];
                        if (stamp[c] != epoch) {
                            stamp[c] = epoch;
                            stack.push_back(c);
                        }
                    }
                }
                const size_t n = part.size() - before;
                if (n <= 64) {
                    std::sort(part.begin() + (ptrdiff_t)before, part.end());
                } else {
                    uint32_t lo = N, hi = 0;
                    for (size_t i = before; i < part.size(); i++) {
                        bits[part[i] >> 6] |= 1ULL << (part[i] & 63);
                        lo = std::min(lo, part[i]);
                        hi = std::max(hi, part[i]);
                    }
                    size_t o = before;
                    for (uint32_t w = lo >> 6; w <= (hi >> 6); w++) {
                        for (uint64_t x = bits[w]; x; x &= x - 1) part[o++] = w * 64 + (uint32_t)__builtin_ctzll(x);
                        bits[w] = 0;
                    }
                }
            }
            counts[t].push_back(part.size() - before);
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; t++) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    out.off.assign(1, 0);
    for (uint32_t t = 0; t < threads; t++)
        for (uint64_t c : counts[t]) out.off.push_back(out.off.back() + c);
    return out;
}

} // namespace grgl_amd
