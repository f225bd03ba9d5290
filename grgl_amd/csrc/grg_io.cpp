// Reader and writer for .grg files, format 5.x.
//
// Layout (src/serialize.cpp:518-669 writes it, :697-890 reads it):
//   128-byte header (include/grgl/common.h:185-202)
//   u64 edge byte count, then per node (in node order) the child list, nodes without
//     children contributing no bytes: LEB128 count, LEB128 deltas of the ascending child
//     IDs starting from 0 (csr_storage.h:440-466 with libvbyte's sorted codec)
//   node byte offsets: numNodes + 1 entries, u32 up to version 5.1, u64 from 5.2
//   mutation -> node (u32 each), 24-byte mutation records (mutation.h:248-261), then the
//     allele strings of records whose allele storage does not fit inline (LEB128 length +
//     bytes)
//   populations (LEB128-length strings) and per-sample u16 population IDs, if any
//   individual coalescence counts: u32 per non-sample node
//   individual IDs (u64 byte count, bytes, CSR offsets) if header flag 0x1
//   missingness node per mutation (u32) if header flag 0x4
#include <algorithm>
#include <cstring>
#include <exception>
#include <fstream>
#include <stdexcept>
#include <thread>

#include "host_graph.h"

namespace grgl_amd {

namespace {

constexpr uint64_t kMagic = 0xE9366C64DDC8C5B0ULL;   // common.h:181
constexpr uint16_t kMajor = 5, kMinor = 3;          // common.h:182-183
constexpr uint64_t kFlagIndivIds = 0x1, kFlagUnphased = 0x2, kFlagMissing = 0x4;

#pragma pack(push, 1)
struct FileHeader {
    uint64_t magic;
    uint16_t versionMajor;
    uint16_t versionMinor;
    uint16_t ploidy;
    uint16_t populationCount;
    uint64_t sampleCount;
    uint64_t mutationCount;
    uint64_t nodeCount;
    uint64_t edgeCount;
    uint64_t rangeStart;
    uint64_t rangeEnd;
    uint64_t flags;
    uint64_t unused[7];
};
struct MutationRecord {
    uint8_t allele[8]; // bit 0 of byte 0 set: the storage string follows the records
    uint64_t position;
    uint32_t time;
    uint16_t altIndex;
    uint16_t unused;
};
#pragma pack(pop)
static_assert(sizeof(FileHeader) == 128, "header size");
static_assert(sizeof(MutationRecord) == 24, "mutation record size");

void fail(const std::string& m) { throw std::runtime_error("Malformed GRG file: " + m); }

// Every array and string is checked against the bytes left in the file before anything is
// allocated for it, so a corrupt count fails as a malformed file instead of a huge allocation.
struct Reader {
    std::ifstream in;
    uint64_t size = 0;
    explicit Reader(const std::string& path) : in(path, std::ios::binary) {
        if (!in) throw std::runtime_error("Could not open GRG file " + path);
        in.seekg(0, std::ios::end);
        const std::streamoff end = in.tellg();
        in.seekg(0, std::ios::beg);
        size = end > 0 ? (uint64_t)end : 0;
    }
    uint64_t remaining() {
        const std::streamoff at = in.tellg();
        return (at >= 0 && (uint64_t)at <= size) ? size - (uint64_t)at : 0;
    }
    void bytes(void* dst, size_t n) {
        in.read(reinterpret_cast<char*>(dst), (std::streamsize)n);
        if ((size_t)in.gcount() != n) fail("unexpected end of file");
    }
    template <typename T> T scalar() {
        T v{};
        bytes(&v, sizeof(T));
        return v;
    }
    template <typename T> std::vector<T> array(uint64_t n) {
        if (n > remaining() / sizeof(T)) fail("unexpected end of file");
        std::vector<T> v(n);
        if (n) bytes(v.data(), n * sizeof(T));
        return v;
    }
    uint32_t varint() {
        uint32_t r = 0;
        for (int shift = 0; shift <= 28; shift += 7) {
            const uint8_t b = scalar<uint8_t>();
            r |= (uint32_t)(b & 0x7f) << shift;
            if (!(b & 0x80)) return r;
        }
        fail("bad varint");
        return 0;
    }
    std::string str() {
        const uint32_t n = varint();
        if (n > remaining()) fail("unexpected end of file");
        std::string s(n, '\0');
        if (n) bytes(&s[0], n);
        return s;
    }
};

size_t putVarint(std::vector<uint8_t>& out, uint32_t v) {
    size_t n = 0;
    while (v >= 0x80) {
        out.push_back((uint8_t)(v | 0x80));
        v >>= 7;
        n++;
    }
    out.push_back((uint8_t)v);
    return n + 1;
}

uint32_t getVarint(const uint8_t* p, size_t avail, size_t& used) {
    uint32_t r = 0;
    used = 0;
    for (int shift = 0; shift <= 28; shift += 7) {
        if (used >= avail) fail("edge list overruns its bytes");
        const uint8_t b = p[used++];
        r |= (uint32_t)(b & 0x7f) << shift;
        if (!(b & 0x80)) return r;
    }
    fail("bad varint in edge list");
    return 0;
}

// Runs body(lo, hi) over [0, n) in contiguous ranges on up to 16 threads (one for small n);
// the first exception thrown by any range is rethrown after all threads finish.
template <typename F> void parallelRanges(uint64_t n, F body) {
    unsigned threads = std::min(16U, std::max(1U, std::thread::hardware_concurrency()));
    if (n < (1U << 16)) threads = 1;
    if (threads == 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> pool;
    std::vector<std::exception_ptr> errors(threads);
    for (unsigned t = 0; t < threads; t++) {
        pool.emplace_back([&, t] {
            try {
                body(n * t / threads, n * (t + 1) / threads);
            } catch (...) {
                errors[t] = std::current_exception();
            }
        });
    }
    for (auto& th : pool) th.join();
    for (auto& e : errors)
        if (e) std::rethrow_exception(e);
}

} // namespace

HostGrg readGrgFile(const std::string& path) {
    Reader r(path);
    const FileHeader h = r.scalar<FileHeader>();
    if (h.magic != kMagic) throw std::runtime_error("Invalid file header");
    if (h.versionMajor != kMajor) throw std::runtime_error("Incompatible file major version");
    if (h.ploidy == 0) fail("ploidy was 0");
    if (h.nodeCount >= kInvalidNode || h.sampleCount > h.nodeCount) fail("node counts");
    HostGrg g;
    g.numSamples = (uint32_t)h.sampleCount;
    g.ploidy = h.ploidy;
    g.rangeStart = h.rangeStart;
    g.rangeEnd = h.rangeEnd;
    g.phased = !(h.flags & kFlagUnphased);
    const uint64_t V = h.nodeCount, M = h.mutationCount;
    // Edges.
    const uint64_t edgeBytes = r.scalar<uint64_t>();
    const std::vector<uint8_t> edges = r.array<uint8_t>(edgeBytes);
    std::vector<uint64_t> index;
    if (h.versionMinor <= 1) {
        const std::vector<uint32_t> small = r.array<uint32_t>(V + 1);
        index.assign(small.begin(), small.end());
    } else {
        index = r.array<uint64_t>(V + 1);
    }
    // Two passes over node ranges in parallel (each list is self-contained at its byte
    // offset): list lengths, then the lists decoded into place.
    g.downOff.assign(V + 1, 0);
    parallelRanges(V, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t n = lo; n < hi; n++) {
            const uint64_t b = index[n], e = index[n + 1];
            if (e < b || e > edgeBytes) fail("node index");
            if (e > b) {
                size_t used = 0;
                const uint32_t count = getVarint(edges.data() + b, (size_t)(e - b), used);
                // every child takes at least one byte of the list
                if (count > (e - b) - used) fail("node list length");
                g.downOff[n + 1] = count;
            }
        }
    });
    for (uint64_t n = 0; n < V; n++) g.downOff[n + 1] += g.downOff[n];
    if (g.downOff[V] != h.edgeCount) fail("edge count does not match the header");
    g.downChild.resize(h.edgeCount);
    parallelRanges(V, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t n = lo; n < hi; n++) {
            const uint64_t b = index[n], e = index[n + 1];
            if (e == b) continue;
            size_t used = 0;
            const uint8_t* p = edges.data() + b;
            size_t avail = (size_t)(e - b);
            const uint32_t count = getVarint(p, avail, used);
            p += used;
            avail -= used;
            uint32_t* out = g.downChild.data() + g.downOff[n];
            uint32_t prev = 0;
            for (uint32_t i = 0; i < count; i++) {
                const uint32_t d = getVarint(p, avail, used);
                p += used;
                avail -= used;
                prev += d;
                out[i] = prev;
            }
        }
    });
    // Mutations.
    g.mutNode = r.array<uint32_t>(M);
    const std::vector<MutationRecord> recs = r.array<MutationRecord>(M);
    g.mutAlleles.resize(M);
    g.mutAltIndex.resize(M);
    g.mutTime.resize(M);
    g.mutPos.resize(M);
    std::vector<uint64_t> deferred;
    for (uint64_t m = 0; m < M; m++) {
        const MutationRecord& rec = recs[m];
        g.mutPos[m] = (double)rec.position;
        g.mutAltIndex[m] = rec.altIndex;
        g.mutTime[m] = rec.time;
        if (rec.allele[0] & 1) {
            deferred.push_back(m);
        } else {
            std::string s;
            for (int i = 1; i < 8 && rec.allele[i]; i++) s.push_back((char)rec.allele[i]);
            g.mutAlleles[m] = s;
        }
    }
    for (uint64_t m : deferred) g.mutAlleles[m] = r.str();
    // Populations.
    for (uint16_t i = 0; i < h.populationCount; i++) g.populations.push_back(r.str());
    if (h.populationCount > 0) g.popIds = r.array<uint16_t>(h.sampleCount);
    // Coalescence counts: stored for non-sample nodes; samples read as 0.
    const std::vector<uint32_t> coals = r.array<uint32_t>(V - h.sampleCount);
    g.coal.assign(V, 0);
    bool anySet = false;
    for (uint64_t i = 0; i < coals.size(); i++) {
        g.coal[h.sampleCount + i] = coals[i];
        anySet |= coals[i] != kCoalNotSet;
    }
    if (!anySet && !coals.empty()) g.coal.clear();
    // Individual IDs.
    if (h.flags & kFlagIndivIds) {
        const uint64_t nbytes = r.scalar<uint64_t>();
        const std::vector<uint8_t> chars = r.array<uint8_t>(nbytes);
        const uint64_t nind = h.sampleCount / h.ploidy;
        std::vector<uint64_t> off(nind + 1);
        if (h.versionMinor <= 1) {
            const std::vector<uint32_t> small = r.array<uint32_t>(nind + 1);
            for (uint64_t i = 0; i <= nind; i++) off[i] = small[i];
        } else {
            off = r.array<uint64_t>(nind + 1);
        }
        for (uint64_t i = 0; i < nind; i++) {
            if (off[i + 1] < off[i] || off[i + 1] > nbytes) fail("individual IDs");
            g.individualIds.emplace_back((const char*)chars.data() + off[i], (size_t)(off[i + 1] - off[i]));
        }
    }
    if (h.flags & kFlagMissing) g.mutMiss = r.array<uint32_t>(M);
    g.mutationsOrdered = true;
    validateHostGrg(g);
    return g;
}

// Simplify-on-write (src/serialize.cpp:387-467, RenumberAndWriteVisitor, applied by
// simplifyAndSerialize to every node of an ordered graph in ID order, grg_helpers.h:58-64):
//   * samples keep their IDs;
//   * a non-sample node's children are its down edges, where a child that was dropped is
//     replaced by that child's own (expanded) children, recursively, adding the dropped child's
//     coalescence count to the node's (serialize.cpp:149-177);
//   * the node is dropped ("extraneous") when it then has at most one child, or at most one
//     parent (up edges in the input graph), or exactly two children and two parents — unless
//     it has children and carries a mutation or is a missingness node (:436-446);
//   * kept nodes are renumbered in visiting order, children sorted by new ID; mutations (and
//     missingness nodes) of dropped nodes become unmapped.
HostGrg simplifyGrg(const HostGrg& g) {
    validateHostGrg(g);
    const uint32_t V = g.numNodes(), N = g.numSamples;
    std::vector<uint32_t> numParents(V, 0);
    for (uint64_t i = 0; i < g.numEdges(); i++) numParents[g.downChild[i]]++;
    std::vector<uint8_t> special(V, 0); // carries a mutation or is a missingness node
    for (uint32_t n : g.mutNode)
        if (n != kInvalidNode) special[n] = 1;
    for (uint32_t n : g.mutMiss)
        if (n != kInvalidNode) special[n] = 1;
    const bool coals = g.hasCoals();
    std::vector<uint32_t> newId(V, kInvalidNode);
    for (uint32_t s = 0; s < N; s++) newId[s] = s;
    HostGrg out = g;
    out.downOff.assign(1, 0);
    out.downChild.clear();
    out.coal.clear();
    if (coals) out.coal.assign(g.coal.begin(), g.coal.begin() + N);
    for (uint32_t s = 0; s < N; s++) out.downOff.push_back(0);
    uint32_t counter = N;
    std::vector<uint32_t> children, stack;
    for (uint32_t n = N; n < V; n++) {
        children.clear();
        uint32_t nodeCoals = coals ? g.coal[n] : kCoalNotSet;
        // Depth-first expansion of dropped children, in down-edge order.
        stack.clear();
        for (uint64_t i = g.downOff[n + 1]; i-- > g.downOff[n];) stack.push_back(g.downChild[i]);
        while (!stack.empty()) {
            const uint32_t c = stack.back();
            stack.pop_back();
            if (newId[c] != kInvalidNode) {
                children.push_back(newId[c]);
                continue;
            }
            if (nodeCoals != kCoalNotSet && coals && g.coal[c] != kCoalNotSet) nodeCoals += g.coal[c];
            for (uint64_t i = g.downOff[c + 1]; i-- > g.downOff[c];) stack.push_back(g.downChild[i]);
        }
        const uint32_t parents = numParents[n];
        const bool extraneous = children.size() <= 1 || parents <= 1 || (children.size() == 2 && parents == 2);
        const bool keepMutNode = !children.empty() && special[n];
        if (extraneous && !keepMutNode) continue;
        newId[n] = counter++;
        std::sort(children.begin(), children.end());
        out.downChild.insert(out.downChild.end(), children.begin(), children.end());
        out.downOff.push_back(out.downChild.size());
        if (coals) out.coal.push_back(nodeCoals);
    }
    for (uint32_t& m : out.mutNode)
        if (m != kInvalidNode) m = newId[m];
    for (uint32_t& m : out.mutMiss)
        if (m != kInvalidNode) m = newId[m];
    return out;
}

namespace {

bool mutationLess(const HostGrg& g, uint32_t a, uint32_t b) {
    const double pa = g.mutPos.empty() ? (double)a : g.mutPos[a];
    const double pb = g.mutPos.empty() ? (double)b : g.mutPos[b];
    if (pa != pb) return pa < pb;
    if (g.mutAlleles.empty()) return false;
    const uint16_t ia = g.mutAltIndex.empty() ? 1 : g.mutAltIndex[a];
    const uint16_t ib = g.mutAltIndex.empty() ? 1 : g.mutAltIndex[b];
    const std::string refA = g.mutAlleles[a].substr(0, ia), refB = g.mutAlleles[b].substr(0, ib);
    if (refA != refB) return refA < refB;
    return g.mutAlleles[a].substr(ia) < g.mutAlleles[b].substr(ib);
}

} // namespace

// Mutations in (position, REF, ALT) order (the reference writes a graph whose mutations are
// not marked ordered in that order, grg.h:528-545 with grg.cpp:189-249; files it reads are
// marked ordered, serialize.cpp:757).
HostGrg sortMutationsForWrite(const HostGrg& g) {
    const uint32_t M = g.numMutations();
    std::vector<uint32_t> order(M);
    for (uint32_t m = 0; m < M; m++) order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return mutationLess(g, a, b); });
    HostGrg out = g;
    auto permute = [&](auto& v) {
        if (v.empty()) return;
        auto copy = v;
        for (uint32_t m = 0; m < M; m++) v[m] = copy[order[m]];
    };
    permute(out.mutNode);
    permute(out.mutMiss);
    permute(out.mutPos);
    permute(out.mutAlleles);
    permute(out.mutAltIndex);
    permute(out.mutTime);
    out.mutationsOrdered = true;
    return out;
}

void writeGrgFile(const HostGrg& input, const std::string& path, bool allowSimplify) {
    HostGrg sorted;
    const HostGrg* src = &input;
    if (!input.mutationsOrdered) {
        sorted = sortMutationsForWrite(input);
        src = &sorted;
    }
    HostGrg simplified;
    if (allowSimplify) {
        simplified = simplifyGrg(*src);
        src = &simplified;
    }
    const HostGrg& g = *src;
    validateHostGrg(g);
    const uint32_t V = g.numNodes(), M = g.numMutations();
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) throw std::runtime_error("Could not open " + path + " for writing");
    FileHeader h{};
    h.magic = kMagic;
    h.versionMajor = kMajor;
    h.versionMinor = kMinor;
    h.ploidy = (uint16_t)g.ploidy;
    h.populationCount = (uint16_t)g.populations.size();
    h.sampleCount = g.numSamples;
    h.mutationCount = M;
    h.nodeCount = V;
    h.edgeCount = g.numEdges();
    h.rangeStart = g.rangeStart;
    h.rangeEnd = g.rangeEnd;
    if (h.rangeEnd == 0 && !g.mutPos.empty()) {
        double hi = 0;
        for (double p : g.mutPos) hi = std::max(hi, p);
        h.rangeEnd = (uint64_t)hi + 1;
    }
    h.flags = (g.phased ? 0 : kFlagUnphased) | (g.mutMiss.empty() ? 0 : kFlagMissing) |
              (g.individualIds.empty() ? 0 : kFlagIndivIds);
    out.write(reinterpret_cast<const char*>(&h), sizeof(h));
    std::vector<uint8_t> edges;
    std::vector<uint64_t> index(V + 1, 0);
    for (uint32_t n = 0; n < V; n++) {
        index[n] = edges.size();
        const uint64_t b = g.downOff[n], e = g.downOff[n + 1];
        if (e > b) {
            putVarint(edges, (uint32_t)(e - b));
            uint32_t prev = 0;
            for (uint64_t i = b; i < e; i++) {
                putVarint(edges, g.downChild[i] - prev);
                prev = g.downChild[i];
            }
        }
    }
    index[V] = edges.size();
    const uint64_t edgeBytes = edges.size();
    out.write(reinterpret_cast<const char*>(&edgeBytes), 8);
    out.write(reinterpret_cast<const char*>(edges.data()), (std::streamsize)edges.size());
    out.write(reinterpret_cast<const char*>(index.data()), (std::streamsize)(index.size() * 8));
    out.write(reinterpret_cast<const char*>(g.mutNode.data()), (std::streamsize)(M * 4));
    std::vector<uint32_t> deferred;
    for (uint32_t m = 0; m < M; m++) {
        MutationRecord rec{};
        const std::string alleles = g.mutAlleles.empty() ? std::string("AG") : g.mutAlleles[m];
        rec.position = g.mutPos.empty() ? m : (uint64_t)g.mutPos[m];
        rec.time = g.mutTime.empty() ? 0x3FFFFFFFU : g.mutTime[m];
        rec.altIndex = g.mutAltIndex.empty() ? 1 : g.mutAltIndex[m];
        if (alleles.size() > 7) {
            rec.allele[0] = 1;
            deferred.push_back(m);
        } else {
            std::memcpy(rec.allele + 1, alleles.data(), alleles.size());
        }
        out.write(reinterpret_cast<const char*>(&rec), sizeof(rec));
    }
    auto writeStr = [&](const std::string& s) {
        std::vector<uint8_t> len;
        putVarint(len, (uint32_t)s.size());
        out.write(reinterpret_cast<const char*>(len.data()), (std::streamsize)len.size());
        out.write(s.data(), (std::streamsize)s.size());
    };
    for (uint32_t m : deferred) writeStr(g.mutAlleles[m]);
    for (const auto& p : g.populations) writeStr(p);
    if (!g.populations.empty()) {
        std::vector<uint16_t> pops = g.popIds;
        pops.resize(g.numSamples, 0xffff);
        out.write(reinterpret_cast<const char*>(pops.data()), (std::streamsize)(pops.size() * 2));
    }
    for (uint32_t n = g.numSamples; n < V; n++) {
        const uint32_t c = g.coal.empty() ? kCoalNotSet : g.coal[n];
        out.write(reinterpret_cast<const char*>(&c), 4);
    }
    if (!g.individualIds.empty()) {
        uint64_t total = 0;
        std::vector<uint64_t> off{0};
        for (const auto& s : g.individualIds) off.push_back(total += s.size());
        out.write(reinterpret_cast<const char*>(&total), 8);
        for (const auto& s : g.individualIds) out.write(s.data(), (std::streamsize)s.size());
        out.write(reinterpret_cast<const char*>(off.data()), (std::streamsize)(off.size() * 8));
    }
    if (!g.mutMiss.empty()) out.write(reinterpret_cast<const char*>(g.mutMiss.data()), (std::streamsize)(M * 4));
    if (!out) throw std::runtime_error("Writing GRG failed");
}

} // namespace grgl_amd
