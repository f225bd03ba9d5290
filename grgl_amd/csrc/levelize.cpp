#include "levelize.h"

#include <algorithm>
#include <cstring>
#include <exception>
#include <fstream>
#include <numeric>
#include <stdexcept>
#include <thread>

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
    // Depth from the roots: a node's row needs every parent's row first.  Nodes without
    // children (the samples, mostly) are read by no other row, so they all go to one last
    // level after every other node: the nodes above them then form many small connected
    // groups (a sample joins the trees of all its parents), which the grouped launches
    // traverse level after level inside one block, and the samples make one large level.
    // A row's gather list does not depend on its level, so neither do its bits.
    std::vector<uint32_t> d(V, 0);
    uint32_t maxInner = 0;
    bool anyInner = false;
    for (uint32_t v = V; v-- > 0;) {
        for (uint64_t i = g.downOff[v]; i < g.downOff[v + 1]; i++) {
            const uint32_t c = g.downChild[i];
            d[c] = std::max(d[c], d[v] + 1);
        }
        if (g.downOff[v + 1] > g.downOff[v]) {
            maxInner = std::max(maxInner, d[v]);
            anyInner = true;
        }
    }
    const uint32_t maxD = anyInner ? maxInner + 1 : 0;
    for (uint32_t v = 0; v < V; v++)
        if (g.downOff[v + 1] == g.downOff[v]) d[v] = maxD;
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

LevelizedGraph levelize(const HostGrg& g) {
    validateHostGrg(g);
    LevelizedGraph lg;
    lg.N = g.numSamples;
    lg.V = g.numNodes();
    lg.M = g.numMutations();
    lg.P = g.ploidy;
    lg.E = g.numEdges();
    lg.hasMiss = !g.mutMiss.empty();
    lg.hasCoal = g.hasCoals();
    for (uint32_t v = lg.N; lg.hasCoal && v < lg.V; v++) lg.hasCoal = g.coal[v] != kCoalNotSet;
    std::exception_ptr downError;
    std::thread downThread([&] {
        try {
            lg.down = buildDownPlan(g);
        } catch (...) {
            downError = std::current_exception();
        }
    });
    try {
        lg.up = buildUpPlan(g);
    } catch (...) {
        downThread.join();
        throw;
    }
    downThread.join();
    if (downError) std::rethrow_exception(downError);
    return lg;
}

namespace {

constexpr char kLevelizedMagic[8] = {'G', 'R', 'G', 'L', 'L', 'V', 'L', 'Z'};
constexpr uint32_t kLevelizedVersion = 1;

// 64-bit checksum over 8-byte words (tail bytes zero-padded): multiply-xorshift mixing.
struct Checksum {
    uint64_t h = 0x243f6a8885a308d3ULL;
    uint64_t pending = 0;
    unsigned npend = 0;
    void word(uint64_t w) {
        h ^= w + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
        h *= 0xbf58476d1ce4e5b9ULL;
        h ^= h >> 31;
    }
    void bytes(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        size_t i = 0;
        while (npend != 0 && i < n) {
            pending |= (uint64_t)b[i++] << (8 * npend);
            if (++npend == 8) {
                word(pending);
                pending = 0;
                npend = 0;
            }
        }
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, b + i, 8);
            word(w);
        }
        for (; i < n; i++) {
            pending |= (uint64_t)b[i] << (8 * npend);
            npend++;
        }
    }
    uint64_t final() {
        if (npend) word(pending);
        word(0x5bd1e995ULL);
        return h;
    }
};

class Writer {
public:
    explicit Writer(const std::string& path) : f_(path, std::ios::binary | std::ios::trunc) {
        if (!f_) throw std::runtime_error("cannot open " + path + " for writing");
    }
    void raw(const void* p, size_t n) {
        sum_.bytes(p, n);
        f_.write(static_cast<const char*>(p), (std::streamsize)n);
    }
    template <typename T> void pod(T v) { raw(&v, sizeof v); }
    void vec(const std::vector<uint32_t>& v) {
        pod<uint64_t>(v.size());
        if (!v.empty()) raw(v.data(), v.size() * sizeof(uint32_t));
    }
    void finish(const std::string& path) {
        const uint64_t c = sum_.final();
        f_.write(reinterpret_cast<const char*>(&c), sizeof c);
        f_.flush();
        if (!f_) throw std::runtime_error("error writing " + path);
    }

private:
    std::ofstream f_;
    Checksum sum_;
};

class Reader {
public:
    explicit Reader(const std::string& path) : path_(path), f_(path, std::ios::binary) {
        if (!f_) throw std::runtime_error("cannot open " + path);
        f_.seekg(0, std::ios::end);
        size_ = (uint64_t)f_.tellg();
        f_.seekg(0, std::ios::beg);
    }
    void raw(void* p, size_t n) {
        if (pos_ + n + sizeof(uint64_t) > size_) bad("truncated file");
        f_.read(static_cast<char*>(p), (std::streamsize)n);
        if (!f_) bad("read error");
        sum_.bytes(p, n);
        pos_ += n;
    }
    template <typename T> T pod() {
        T v;
        raw(&v, sizeof v);
        return v;
    }
    void vec(std::vector<uint32_t>& v, uint64_t expect) {
        const uint64_t n = pod<uint64_t>();
        if (n != expect) bad("array length does not match the graph");
        if (n > (size_ - pos_) / sizeof(uint32_t)) bad("truncated file");
        v.resize(n);
        if (n) raw(v.data(), n * sizeof(uint32_t));
    }
    void finish() {
        uint64_t c = 0;
        if (pos_ + sizeof c != size_) bad("trailing bytes");
        f_.read(reinterpret_cast<char*>(&c), sizeof c);
        if (!f_ || c != sum_.final()) bad("checksum mismatch (corrupt file)");
    }
    [[noreturn]] void bad(const std::string& why) const {
        throw std::runtime_error("not a valid levelized graph file (" + why + "): " + path_);
    }

private:
    std::string path_;
    std::ifstream f_;
    uint64_t size_ = 0, pos_ = 0;
    Checksum sum_;
};

void writePlan(Writer& w, const DirPlan& p) {
    w.pod<uint32_t>(p.numNodes);
    w.pod<uint32_t>(p.numLevels);
    w.pod<uint32_t>(p.firstPulledLevel);
    w.pod<uint32_t>(p.maxFanIn);
    w.pod<uint64_t>(p.src.size());
    w.vec(p.levelStart);
    w.vec(p.order);
    w.vec(p.rank);
    w.vec(p.off);
    w.vec(p.src);
    w.vec(p.outRow);
    w.vec(p.missRow);
    w.vec(p.coalByRow);
}

void readPlan(Reader& r, DirPlan& p, const LevelizedGraph& lg, bool up) {
    p.numNodes = r.pod<uint32_t>();
    p.numLevels = r.pod<uint32_t>();
    p.firstPulledLevel = r.pod<uint32_t>();
    p.maxFanIn = r.pod<uint32_t>();
    const uint64_t entries = r.pod<uint64_t>();
    if (p.numNodes != lg.V) r.bad("plan size does not match the graph");
    if (p.numLevels == 0 || p.numLevels > lg.V + 1) r.bad("bad level count");
    r.vec(p.levelStart, (uint64_t)p.numLevels + 1);
    r.vec(p.order, lg.V);
    r.vec(p.rank, lg.V);
    r.vec(p.off, (uint64_t)lg.V + 1);
    r.vec(p.src, entries);
    r.vec(p.outRow, up ? lg.M : lg.N);
    r.vec(p.missRow, (up && lg.hasMiss) ? lg.M : 0);
    r.vec(p.coalByRow, lg.V);
}

void validatePlan(const DirPlan& p, const LevelizedGraph& lg, bool up) {
    auto bad = [&](const char* why) {
        throw std::runtime_error(std::string("invalid levelized graph (") + (up ? "UP" : "DOWN") + " plan): " + why);
    };
    const uint32_t V = lg.V;
    if (p.numNodes != V || p.order.size() != V || p.rank.size() != V || p.off.size() != (size_t)V + 1 ||
        p.levelStart.size() != (size_t)p.numLevels + 1 || p.coalByRow.size() != V)
        bad("array sizes");
    if (p.firstPulledLevel != (up ? 1U : 0U)) bad("first pulled level");
    if (p.levelStart[0] != 0 || p.levelStart[p.numLevels] != V) bad("level bounds");
    for (uint32_t l = 0; l < p.numLevels; l++)
        if (p.levelStart[l + 1] < p.levelStart[l]) bad("level bounds");
    for (uint32_t j = 0; j < V; j++) {
        if (p.order[j] >= V || p.rank[p.order[j]] != j) bad("row order is not a permutation");
    }
    if (up) {
        if (p.levelStart[1] < lg.N) bad("samples outside level 0");
        for (uint32_t s = 0; s < lg.N; s++)
            if (p.order[s] != s) bad("samples not first in sample order");
    }
    if (p.off[0] != 0) bad("list offsets");
    // Seed rows (DOWN): V + m, and V + M + m with missing data.
    const uint64_t seedEnd = up ? V : (uint64_t)V + (uint64_t)lg.M * (lg.hasMiss ? 2 : 1);
    uint64_t pos = 0;
    uint32_t maxLen = 0;
    for (uint32_t l = 0; l < p.numLevels; l++) {
        uint32_t prev = 0xffffffffU;
        for (uint32_t j = p.levelStart[l]; j < p.levelStart[l + 1]; j++) {
            const uint32_t len = p.off[j + 1] - p.off[j]; // modulo 2^32, as the kernels read it
            if (len >= (1U << 31)) bad("gather list too long");
            if (l < p.firstPulledLevel && len != 0) bad("seeded rows with gather lists");
            if (l >= p.firstPulledLevel && len > prev) bad("gather lists not longest first within a level");
            prev = len;
            maxLen = std::max(maxLen, len);
            if (pos + len > p.src.size()) bad("gather lists overrun the entries");
            for (uint64_t e = pos; e < pos + len; e++) {
                const uint32_t s = p.src[e];
                const bool earlier = s < p.levelStart[l];
                const bool seed = !up && s >= V && (uint64_t)s < seedEnd;
                if (!earlier && !seed) bad("gather entry outside the earlier levels");
            }
            pos += len;
        }
    }
    if (pos != p.src.size()) bad("entry count");
    if (maxLen != p.maxFanIn) bad("maximum fan-in");
    for (uint32_t x : p.outRow)
        if (x >= V && !(up && x == kInvalidNode)) bad("read-out row out of range");
    for (uint32_t x : p.missRow)
        if (x >= V && x != kInvalidNode) bad("missingness row out of range");
    for (uint32_t j = 0; j < V; j++) {
        const bool sample = p.order[j] < lg.N;
        if (sample && p.coalByRow[j] != 0) bad("sample coalescence counts");
        if (!sample && lg.hasCoal && p.coalByRow[j] == kCoalNotSet) bad("missing coalescence counts");
    }
}

} // namespace

void validateLevelized(const LevelizedGraph& lg) {
    if (lg.N == 0 || lg.P == 0 || lg.N % lg.P != 0 || lg.V < lg.N || lg.V >= kInvalidNode)
        throw std::runtime_error("invalid levelized graph: sizes");
    if ((uint64_t)lg.V + 2ULL * lg.M >= 0x7fffffffULL)
        throw std::runtime_error("invalid levelized graph: nodes + 2*mutations must stay below 2^31");
    if (lg.up.src.size() != lg.E) throw std::runtime_error("invalid levelized graph: UP entries != edges");
    validatePlan(lg.up, lg, true);
    validatePlan(lg.down, lg, false);
    if (!lg.down.missRow.empty()) throw std::runtime_error("invalid levelized graph: DOWN missingness rows");
}

void writeLevelizedFile(const LevelizedGraph& lg, const std::string& path) {
    Writer w(path);
    w.raw(kLevelizedMagic, sizeof kLevelizedMagic);
    w.pod<uint32_t>(kLevelizedVersion);
    w.pod<uint32_t>(lg.N);
    w.pod<uint32_t>(lg.V);
    w.pod<uint32_t>(lg.M);
    w.pod<uint32_t>(lg.P);
    w.pod<uint32_t>((lg.hasMiss ? 1U : 0U) | (lg.hasCoal ? 2U : 0U));
    w.pod<uint64_t>(lg.E);
    writePlan(w, lg.up);
    writePlan(w, lg.down);
    w.finish(path);
}

LevelizedGraph readLevelizedFile(const std::string& path) {
    Reader r(path);
    char magic[8];
    r.raw(magic, sizeof magic);
    if (std::memcmp(magic, kLevelizedMagic, sizeof magic) != 0) r.bad("magic");
    if (r.pod<uint32_t>() != kLevelizedVersion) r.bad("unsupported format version");
    LevelizedGraph lg;
    lg.N = r.pod<uint32_t>();
    lg.V = r.pod<uint32_t>();
    lg.M = r.pod<uint32_t>();
    lg.P = r.pod<uint32_t>();
    const uint32_t flags = r.pod<uint32_t>();
    if (flags & ~3U) r.bad("flags");
    lg.hasMiss = (flags & 1U) != 0;
    lg.hasCoal = (flags & 2U) != 0;
    lg.E = r.pod<uint64_t>();
    if (lg.V >= kInvalidNode || lg.N > lg.V) r.bad("sizes");
    readPlan(r, lg.up, lg, true);
    readPlan(r, lg.down, lg, false);
    r.finish();
    validateLevelized(lg);
    return lg;
}

} // namespace grgl_amd
