// Mutable GRG and mutation mapping (see construct.h).
//
// Mapping one mutation with sample set S onto the graph.  This is the reference's greedy
// candidate cover (mapMutations, src/map_mutations.cpp:463-607 for the candidates, :609-728
// for the cover, :758-876 for batches), found by a different sweep:
//   1. Coverage sweep.  Every sample of S starts with count 1.  Touched nodes are visited
//      in increasing ID order (children before parents, since IDs are ordered); a node is
//      "covered" when its count equals the number of samples below it, and only covered
//      nodes pass their count to their parents, so a parent reaches its own sample count
//      exactly when all its children are covered.  Nodes created in the current batch are
//      not reuse candidates.
//   2. A covered node with |S| samples is reused as the mutation's node.
//   3. Otherwise the maximal covered nodes (no covered parent) of at least two samples are
//      taken largest first, each only if its samples are disjoint from those already taken
//      (two candidates can share samples in a DAG), and S's remaining samples become direct
//      children of a new node.  The reference also lists a covered node with a covered and an
//      uncovered parent; its covered parent is larger and taken first, so the cover is the
//      same.  Equal-size candidates go highest node ID first (the reference's order among
//      them is unspecified: std::sort by size).  Children of a node therefore have disjoint
//      sample sets: every sample reaches every node by at most one path, which keeps matmul
//      counts exact.
//   4. For diploid graphs the new node records how many individuals have their two
//      haplotypes under different children (the coalescence count of grg_from_trees,
//      src/python/_grgl.cpp:1018-1021).
#include "construct.h"

#include <algorithm>
#include <functional>
#include <queue>
#include <stdexcept>

namespace grgl_amd {

namespace {
[[noreturn]] void misuse(const std::string& m) { throw std::runtime_error(m); }
constexpr uint32_t kPloidyCoal = 2;
} // namespace

MutableGraph::MutableGraph(uint32_t numSamples, uint32_t ploidy, bool phased)
    : numSamples_(numSamples), ploidy_(ploidy), phased_(phased) {
    if (ploidy == 0) misuse("ploidy must be at least 1");
    if (numSamples % ploidy != 0) misuse("num_samples must be a multiple of ploidy");
    down_.resize(numSamples);
    up_.resize(numSamples);
    count_.assign(numSamples, 1);
    coal_.assign(numSamples, kCoalNotSet);
}

uint32_t MutableGraph::makeNode(uint32_t count) {
    const uint64_t first = down_.size();
    if (first + count >= (uint64_t)kInvalidNode) misuse("too many nodes");
    down_.resize(first + count);
    up_.resize(first + count);
    count_.resize(first + count, 0);
    coal_.resize(first + count, kCoalNotSet);
    version_++;
    return (uint32_t)first;
}

void MutableGraph::connect(uint32_t parent, uint32_t child) {
    if (parent >= numNodes() || child >= numNodes()) misuse("connect: node ID out of range");
    if (child >= parent) misuse("connect: the child must have a smaller node ID than the parent (ordered nodes)");
    if (parent < numSamples_) misuse("connect: sample nodes cannot have children");
    down_[parent].push_back(child);
    up_[child].push_back(parent);
    count_[parent] += count_[child];
    numEdges_++;
    version_++;
}

uint32_t MutableGraph::addMutation(const MutationRecordIn& m, uint32_t node, uint32_t missNode) {
    if (node != kInvalidNode && node >= numNodes()) misuse("add_mutation: node ID out of range");
    if (missNode != kInvalidNode && missNode >= numNodes()) misuse("add_mutation: missingness node out of range");
    if (m.refAllele.size() > 0xffff) misuse("REF allele length exceeded maximum of 65535");
    mutNode_.push_back(node);
    mutMiss_.push_back(missNode);
    anyMiss_ = anyMiss_ || missNode != kInvalidNode;
    mutPos_.push_back(m.position);
    mutAlleles_.push_back(m.refAllele + m.allele);
    mutAltIndex_.push_back((uint16_t)m.refAllele.size());
    version_++;
    return (uint32_t)(mutNode_.size() - 1);
}

void MutableGraph::setNumIndividualCoals(uint32_t node, uint32_t coals) {
    if (node >= numNodes()) misuse("node ID out of range");
    coal_[node] = coals;
    anyCoal_ = true;
    version_++;
}

MappingStats MutableGraph::mapMutations(const std::vector<MutationRecordIn>& muts,
                                        const std::vector<std::vector<uint32_t>>& samples, size_t batchSize,
                                        bool sizeOrder) {
    if (muts.size() != samples.size()) misuse("mutations and samples must be the same length in mapMutations");
    // Sample counts from scratch: connect() keeps them incrementally, which is stale when a
    // node gained children after it was connected to its parents.
    for (uint32_t n = numSamples_; n < numNodes(); n++) {
        uint32_t c = 0;
        for (uint32_t ch : down_[n]) c += count_[ch];
        count_[n] = c;
    }
    MappingStats st;
    st.totalMutations = muts.size();
    const size_t n = muts.size();
    std::vector<std::vector<uint32_t>> sets(n);
    for (size_t i = 0; i < n; i++) {
        sets[i] = samples[i];
        std::sort(sets[i].begin(), sets[i].end());
        sets[i].erase(std::unique(sets[i].begin(), sets[i].end()), sets[i].end());
        for (uint32_t x : sets[i])
            if (x >= numSamples_) misuse("map_mutations: sample ID out of range");
        if (sets[i].empty() && muts[i].allele == ".") misuse("Missing data rows cannot have no samples");
    }
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    if (sizeOrder) {
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return sets[a].size() < sets[b].size(); });
    }
    if (batchSize == 0) batchSize = n ? n : 1;
    std::vector<uint32_t> nodeOf(n, kInvalidNode);
    uint32_t limit = numNodes();
    for (size_t k = 0; k < n; k++) {
        if (k % batchSize == 0) limit = numNodes(); // new batch: nodes made so far become candidates
        nodeOf[order[k]] = mapSet(sets[order[k]], limit, st);
    }
    // Mutations are recorded in input order.  Missing-data rows ("." alleles) precede the
    // other rows of their site and give those rows their missingness node.
    for (size_t i = 0; i < n; i++) {
        const MutationRecordIn& m = muts[i];
        const uint32_t missNode = (missPos_ == m.position) ? missNode_ : kInvalidNode;
        if (m.allele == ".") {
            missPos_ = m.position;
            missNode_ = nodeOf[i];
        } else {
            addMutation(m, nodeOf[i], missNode);
        }
    }
    return st;
}

uint32_t MutableGraph::mapSet(const std::vector<uint32_t>& S, uint32_t limit, MappingStats& st) {
    st.samplesProcessed += S.size();
    if (S.empty()) {
        st.emptyMutations++;
        return kInvalidNode;
    }
    if (S.size() == 1) {
        st.mutationsWithOneSample++;
        st.mutationsWithNoCandidates++;
        if (ploidy_ == kPloidyCoal) setNumIndividualCoals(S[0], 0);
        return S[0];
    }

    // 1. Coverage sweep over the touched part of the graph.
    const uint32_t V = numNodes();
    if (cnt_.size() < V) {
        cnt_.resize(V, 0);
        stamp_.resize(V, 0);
    }
    if (sampleMark_.size() < numSamples_) {
        sampleMark_.assign(numSamples_, 0);
        ownerChild_.assign(numSamples_, kInvalidNode);
    }
    if (epoch_ >= 0x7ffffff0U) {
        std::fill(stamp_.begin(), stamp_.end(), 0);
        std::fill(sampleMark_.begin(), sampleMark_.end(), 0);
        epoch_ = 0;
    }
    const uint32_t touched = ++epoch_ * 2;  // stamp value: touched this mutation
    const uint32_t covered = touched + 1;   // stamp value: touched and covered
    std::priority_queue<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> heap;
    for (uint32_t s : S) {
        cnt_[s] = 1;
        stamp_[s] = touched;
        heap.push(s);
    }
    std::vector<uint32_t> coveredNodes;
    while (!heap.empty()) {
        const uint32_t v = heap.top();
        heap.pop();
        if (cnt_[v] != count_[v]) continue;
        stamp_[v] = covered;
        if (v >= numSamples_) coveredNodes.push_back(v);
        for (uint32_t p : up_[v]) {
            if (p >= limit) continue;
            if (stamp_[p] < touched) {
                stamp_[p] = touched;
                cnt_[p] = 0;
                heap.push(p);
            }
            cnt_[p] += cnt_[v];
        }
    }
    std::vector<uint32_t> cands;
    for (uint32_t v : coveredNodes) {
        if (count_[v] < 2) continue; // candidates cover at least two samples (map_mutations.cpp:512)
        bool maximal = true;
        for (uint32_t p : up_[v])
            if (p < limit && stamp_[p] == covered) {
                maximal = false;
                break;
            }
        if (maximal) cands.push_back(v);
    }
    st.numCandidates += cands.size();
    if (cands.empty()) st.mutationsWithNoCandidates++;

    // 2. Exact reuse.
    uint32_t exact = kInvalidNode;
    for (uint32_t v : cands)
        if (count_[v] == S.size() && (exact == kInvalidNode || v > exact)) exact = v;
    if (exact != kInvalidNode) {
        st.reusedExactly++;
        if (exact >= limit) st.reusedMutNodes++;
        return exact;
    }

    // 3. Greedy disjoint cover by maximal candidates, largest first.
    std::sort(cands.begin(), cands.end(), [&](uint32_t a, uint32_t b) {
        return count_[a] != count_[b] ? count_[a] > count_[b] : a > b;
    });
    const uint32_t mark = epoch_;
    std::vector<uint32_t> children, below, stack;
    std::vector<uint32_t> visit; // nodes visited while listing one candidate's samples
    uint64_t nCovered = 0;
    for (uint32_t c : cands) {
        if (nCovered == S.size()) break;
        below.clear();
        stack.assign(1, c);
        visit.clear();
        // List the candidate's samples (each is reached once: children are disjoint).
        while (!stack.empty()) {
            const uint32_t v = stack.back();
            stack.pop_back();
            if (v < numSamples_) {
                below.push_back(v);
                continue;
            }
            for (uint32_t ch : down_[v]) stack.push_back(ch);
        }
        bool overlaps = false;
        for (uint32_t s : below)
            if (sampleMark_[s] == mark) {
                overlaps = true;
                break;
            }
        if (overlaps) continue;
        for (uint32_t s : below) {
            sampleMark_[s] = mark;
            ownerChild_[s] = c;
        }
        nCovered += below.size();
        children.push_back(c);
        st.reusedNodes++;
        st.reusedNodeCoverage += below.size();
        if (below.size() < st.reuseSizeHist.size()) st.reuseSizeHist[below.size()]++;
        else st.reuseSizeBiggerThanHistMax++;
        if (c >= limit) st.reusedMutNodes++;
    }
    uint64_t singles = 0;
    for (uint32_t s : S) {
        if (sampleMark_[s] != mark) {
            ownerChild_[s] = s;
            children.push_back(s);
            singles++;
        }
    }
    if (singles) st.numWithSingletons++;
    st.maxSingletons = std::max<uint64_t>(st.maxSingletons, singles);
    st.singletonSampleEdges += singles;

    const uint32_t node = makeNode(1);
    std::sort(children.begin(), children.end());
    for (uint32_t c : children) connect(node, c);
    st.newTreeNodes++;
    if (ploidy_ == kPloidyCoal) {
        uint32_t coals = 0;
        for (size_t i = 0; i + 1 < S.size(); i++) {
            if ((S[i] & 1U) == 0 && S[i + 1] == S[i] + 1 && ownerChild_[S[i]] != ownerChild_[S[i + 1]]) coals++;
        }
        setNumIndividualCoals(node, coals);
    }
    return node;
}

HostGrg MutableGraph::freeze() const {
    HostGrg g;
    g.numSamples = numSamples_;
    g.ploidy = ploidy_;
    g.phased = phased_;
    const uint32_t V = numNodes();
    g.downOff.assign((size_t)V + 1, 0);
    for (uint32_t n = 0; n < V; n++) g.downOff[n + 1] = g.downOff[n] + down_[n].size();
    g.downChild.reserve(g.downOff[V]);
    for (uint32_t n = 0; n < V; n++) {
        std::vector<uint32_t> k(down_[n]);
        std::sort(k.begin(), k.end());
        g.downChild.insert(g.downChild.end(), k.begin(), k.end());
    }
    g.mutNode = mutNode_;
    if (anyMiss_) g.mutMiss = mutMiss_;
    g.mutPos = mutPos_;
    g.mutAlleles = mutAlleles_;
    g.mutAltIndex = mutAltIndex_;
    if (anyCoal_) g.coal = coal_;
    // Adding mutations to a mutable graph leaves them unordered (grg.cpp:104-106): the .grg
    // writer sorts them.
    g.mutationsOrdered = g.mutNode.empty();
    validateHostGrg(g);
    return g;
}

} // namespace grgl_amd
