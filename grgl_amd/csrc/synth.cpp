// Seeded synthetic GRGs from Kingman coalescent trees, plus HostGrg validation and stats.
//
// Per tree (one per genome segment):
//   1. n lineages merge pairwise, the pair chosen uniformly, after Exp(rate = k(k-1)/2)
//      waiting times (k = lineages left).  Internal node IDs grow with merge time, so a
//      parent's ID is always larger than its children's.
//   2. Each of the tree's sites gets one mutation on a branch chosen with probability
//      proportional to branch length, at a uniform position inside the segment.
//   3. GRG nodes follow the reference's simplify-on-write rule (src/serialize.cpp:436-463,
//      RenumberAndWriteVisitor::visit): every internal tree node has one parent, so it is
//      kept only when it carries a mutation; a dropped node's children move up to its
//      nearest kept ancestor (getChildren, serialize.cpp:149-172).  A mutation on a leaf
//      branch maps to the sample node itself.  Kept nodes are renumbered in depth-first
//      post-order, samples first, as the reference's writer does.
//   4. Individual coalescence counts (ploidy 2): an individual coalesces at the tree node
//      that is the MRCA of its two haplotypes; a dropped node's count moves to its nearest
//      kept ancestor, as serialize.cpp:155-163 propagates parentCoals.
#include "host_graph.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace grgl_amd {

namespace {

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// xoshiro256** (Blackman & Vigna), seeded through splitmix64.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (auto& w : s) w = splitmix64(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    uint64_t below(uint64_t n) { return (uint64_t)(uniform() * (double)n) % n; }
    double exponential(double rate) { return -std::log1p(-uniform()) / rate; }
};

constexpr uint32_t kNone = 0xffffffffU;

struct TreePart {
    uint32_t keptInternal = 0;
    std::vector<uint32_t> childOff;  // keptInternal + 1
    std::vector<uint32_t> children;  // samples < n, internal as n + localId
    std::vector<double> mutPos;
    std::vector<uint32_t> mutNode;   // sample id, or n + localId
    std::vector<uint32_t> coal;      // per local internal
};

TreePart buildTree(uint32_t n, uint32_t numMuts, double lo, double hi, uint64_t seed, uint32_t ploidy) {
    Rng rng(seed);
    const uint32_t total = 2 * n - 1;
    std::vector<uint32_t> parent(total, kNone);
    std::vector<double> time(total, 0.0);
    std::vector<uint32_t> c0(total, kNone), c1(total, kNone);
    {
        std::vector<uint32_t> active(n);
        std::iota(active.begin(), active.end(), 0U);
        uint32_t k = n, next = n;
        double t = 0.0;
        while (k > 1) {
            t += rng.exponential(0.5 * (double)k * (double)(k - 1));
            uint32_t i = (uint32_t)rng.below(k);
            uint32_t j = (uint32_t)rng.below(k - 1);
            if (j >= i) j++;
            const uint32_t a = active[i], b = active[j];
            parent[a] = parent[b] = next;
            c0[next] = a;
            c1[next] = b;
            time[next] = t;
            const uint32_t lo_i = std::min(i, j), hi_i = std::max(i, j);
            active[lo_i] = next;
            active[hi_i] = active[k - 1];
            k--;
            next++;
        }
    }
    const uint32_t root = total - 1;
    TreePart part;
    // Mutations on branches, probability proportional to branch length.
    std::vector<uint32_t> mutCount(total, 0);
    std::vector<std::pair<double, uint32_t>> muts(numMuts);
    if (n > 1 && numMuts > 0) {
        std::vector<double> cum(total - 1);
        double acc = 0.0;
        for (uint32_t v = 0; v < root; v++) {
            acc += time[parent[v]] - time[v];
            cum[v] = acc;
        }
        for (uint32_t m = 0; m < numMuts; m++) {
            const double u = rng.uniform() * acc;
            uint32_t v = (uint32_t)(std::upper_bound(cum.begin(), cum.end(), u) - cum.begin());
            if (v >= root) v = root - 1;
            muts[m] = {lo + rng.uniform() * (hi - lo), v};
            mutCount[v]++;
        }
        std::sort(muts.begin(), muts.end());
    }
    // Kept nodes and their nearest kept strict ancestor.
    auto kept = [&](uint32_t v) { return v < n || mutCount[v] > 0; };
    std::vector<uint32_t> keptAnc(total, kNone);
    for (uint32_t v = root; v-- > 0;) {
        const uint32_t p = parent[v];
        keptAnc[v] = kept(p) ? p : keptAnc[p];
    }
    // Post-order numbering of kept internal nodes.
    std::vector<uint32_t> localId(total, kNone);
    {
        std::vector<std::pair<uint32_t, uint8_t>> stack;
        stack.reserve(64);
        stack.push_back({root, 0});
        while (!stack.empty()) {
            auto& top = stack.back();
            const uint32_t v = top.first;
            if (v < n) {
                stack.pop_back();
                continue;
            }
            if (top.second == 0) {
                top.second = 1;
                stack.push_back({c0[v], 0});
            } else if (top.second == 1) {
                top.second = 2;
                stack.push_back({c1[v], 0});
            } else {
                stack.pop_back();
                if (kept(v)) localId[v] = part.keptInternal++;
            }
        }
    }
    auto globalOf = [&](uint32_t v) { return v < n ? v : n + localId[v]; };
    // Children lists: each kept node hangs under its nearest kept ancestor.
    const uint32_t K = part.keptInternal;
    part.childOff.assign(K + 1, 0);
    for (uint32_t v = 0; v < root; v++) {
        if (kept(v) && keptAnc[v] != kNone) part.childOff[localId[keptAnc[v]] + 1]++;
    }
    for (uint32_t i = 0; i < K; i++) part.childOff[i + 1] += part.childOff[i];
    part.children.resize(part.childOff[K]);
    {
        std::vector<uint32_t> fill(part.childOff.begin(), part.childOff.end() - 1);
        for (uint32_t v = 0; v < root; v++) {
            if (kept(v) && keptAnc[v] != kNone) part.children[fill[localId[keptAnc[v]]]++] = globalOf(v);
        }
        for (uint32_t i = 0; i < K; i++) {
            std::sort(part.children.begin() + part.childOff[i], part.children.begin() + part.childOff[i + 1]);
        }
    }
    part.mutPos.resize(numMuts);
    part.mutNode.resize(numMuts);
    for (uint32_t m = 0; m < numMuts; m++) {
        part.mutPos[m] = muts[m].first;
        part.mutNode[m] = globalOf(muts[m].second);
    }
    if (ploidy == 2) {
        std::vector<uint32_t> coalAt(total, 0);
        for (uint32_t i = 0; i + 1 < n; i += 2) {
            uint32_t a = i, b = i + 1;
            while (a != b) {
                if (time[a] <= time[b]) a = parent[a];
                else b = parent[b];
            }
            coalAt[a]++;
        }
        part.coal.assign(K, 0);
        for (uint32_t u = n; u < total; u++) {
            if (!coalAt[u]) continue;
            const uint32_t owner = kept(u) ? u : keptAnc[u];
            if (owner != kNone) part.coal[localId[owner]] += coalAt[u];
        }
    }
    return part;
}

} // namespace

HostGrg synthCoalescentGrg(const SynthParams& p) {
    if (p.numSamples < 2) throw std::runtime_error("synthetic GRG needs at least 2 samples");
    if (p.numTrees < 1) throw std::runtime_error("synthetic GRG needs at least 1 tree");
    if (p.ploidy < 1 || p.numSamples % p.ploidy != 0) throw std::runtime_error("numSamples must be a multiple of ploidy");
    const uint32_t T = p.numTrees, n = p.numSamples;
    std::vector<TreePart> parts(T);
    unsigned threads = p.threads ? p.threads : std::max(1U, std::thread::hardware_concurrency());
    threads = std::min<unsigned>(threads, T);
    auto work = [&](unsigned tid) {
        for (uint32_t t = tid; t < T; t += threads) {
            uint64_t s = p.seed ^ (0xd1b54a32d192ed03ULL * (uint64_t)(t + 1));
            const uint64_t treeSeed = splitmix64(s);
            const uint32_t muts = p.numSites / T + (t < p.numSites % T ? 1 : 0);
            const double lo = p.genomeLength * (double)t / (double)T;
            const double hi = p.genomeLength * (double)(t + 1) / (double)T;
            parts[t] = buildTree(n, muts, lo, hi, treeSeed, p.ploidy);
        }
    };
    if (threads <= 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (unsigned i = 0; i < threads; i++) pool.emplace_back(work, i);
        for (auto& th : pool) th.join();
    }
    HostGrg g;
    g.numSamples = n;
    g.ploidy = p.ploidy;
    uint64_t V = n, E = 0, M = 0;
    for (auto& part : parts) {
        V += part.keptInternal;
        E += part.children.size();
        M += part.mutNode.size();
    }
    if (V >= kInvalidNode) throw std::runtime_error("synthetic GRG exceeds the node limit");
    g.downOff.assign(V + 1, 0);
    g.downChild.resize(E);
    g.mutNode.resize(M);
    g.mutPos.resize(M);
    if (p.ploidy == 2) g.coal.assign(V, 0);
    uint64_t base = n, e = 0, m = 0;
    for (auto& part : parts) {
        const uint32_t shift = (uint32_t)(base - n);
        for (uint32_t i = 0; i < part.keptInternal; i++) {
            g.downOff[base + i] = e;
            for (uint32_t j = part.childOff[i]; j < part.childOff[i + 1]; j++) {
                const uint32_t c = part.children[j];
                g.downChild[e++] = c < n ? c : c + shift;
            }
            if (!g.coal.empty()) g.coal[base + i] = part.coal[i];
        }
        for (size_t j = 0; j < part.mutNode.size(); j++) {
            const uint32_t v = part.mutNode[j];
            g.mutNode[m] = v < n ? v : v + shift;
            g.mutPos[m] = part.mutPos[j];
            m++;
        }
        base += part.keptInternal;
        part = TreePart();
    }
    for (uint64_t v = base; v <= V; v++) g.downOff[v] = e;
    // Samples have no children.
    for (uint32_t s = 0; s < n; s++) g.downOff[s] = 0;
    return g;
}

} // namespace grgl_amd
