// Validation and shape statistics of host-side GRGs.
#include "host_graph.h"

#include <algorithm>
#include <stdexcept>

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

} // namespace grgl_amd
