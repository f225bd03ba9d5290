// Mutable GRG and mutation mapping: the construction side of the reference's API
// (MutableGRG, src/python/_grgl.cpp:759-800; mapMutations, src/map_mutations.cpp:758-876).
//
// A MutableGraph only ever grows: new nodes take the next ID and their children must be
// existing nodes, so every child ID is smaller than its parent's (the "ordered nodes"
// property the reference requires before mapping, map_mutations.cpp:779-782).  freeze()
// produces the immutable CSR form (HostGrg) that the device library uploads.
#pragma once

#include "host_graph.h"

#include <cstdint>
#include <string>
#include <vector>

namespace grgl_amd {

struct MutationRecordIn {
    double position = 0;
    std::string allele;    // "." marks a missing-data row (mutation.cpp:9)
    std::string refAllele;
};

// Fields of the reference's MutationMappingStats (include/grgl/map_mutations.h:164-180).
struct MappingStats {
    uint64_t totalMutations = 0, emptyMutations = 0, mutationsWithOneSample = 0, mutationsWithNoCandidates = 0;
    uint64_t reusedNodes = 0, reusedNodeCoverage = 0, reusedExactly = 0, singletonSampleEdges = 0;
    uint64_t newTreeNodes = 0, samplesProcessed = 0, numCandidates = 0, reuseSizeBiggerThanHistMax = 0;
    uint64_t numWithSingletons = 0, maxSingletons = 0, reusedMutNodes = 0;
    std::vector<uint64_t> reuseSizeHist = std::vector<uint64_t>(1000, 0);
};

class MutableGraph {
public:
    MutableGraph(uint32_t numSamples, uint32_t ploidy, bool phased = true);

    uint32_t numSamples() const { return numSamples_; }
    uint32_t ploidy() const { return ploidy_; }
    uint32_t numNodes() const { return (uint32_t)down_.size(); }
    uint64_t numEdges() const { return numEdges_; }
    uint32_t numMutations() const { return (uint32_t)mutNode_.size(); }
    const std::vector<uint32_t>& children(uint32_t n) const { return down_.at(n); }
    const std::vector<uint32_t>& parents(uint32_t n) const { return up_.at(n); }
    uint32_t sampleCount(uint32_t n) const { return count_.at(n); }
    uint64_t version() const { return version_; }

    // Returns the first of `count` new node IDs.
    uint32_t makeNode(uint32_t count = 1);
    // Edge parent -> child; the child must have a smaller ID (ordered nodes).
    void connect(uint32_t parent, uint32_t child);
    // Returns the new MutationId.
    uint32_t addMutation(const MutationRecordIn& m, uint32_t node, uint32_t missNode = kInvalidNode);
    void setNumIndividualCoals(uint32_t node, uint32_t coals);

    // Maps each mutation onto the graph so that the samples below its node are exactly the
    // given samples; batches of `batchSize` mutations see only nodes that existed when the
    // batch started as reuse candidates (0 = one batch).  With sizeOrder the sets are mapped
    // smallest first (so small sets become building blocks of larger ones); mutation IDs
    // follow the input order either way.
    MappingStats mapMutations(const std::vector<MutationRecordIn>& muts,
                              const std::vector<std::vector<uint32_t>>& samples, size_t batchSize,
                              bool sizeOrder = false);

    HostGrg freeze() const;

private:
    // Node whose samples are exactly S (sorted, unique), reusing or creating nodes.
    uint32_t mapSet(const std::vector<uint32_t>& S, uint32_t candidateLimit, MappingStats& st);

    uint32_t numSamples_, ploidy_;
    bool phased_;
    std::vector<std::vector<uint32_t>> down_, up_;
    std::vector<uint32_t> count_; // samples below each node (each reached by exactly one path)
    std::vector<uint32_t> coal_;
    bool anyCoal_ = false;
    uint64_t numEdges_ = 0;
    uint64_t version_ = 0;
    std::vector<uint32_t> mutNode_, mutMiss_;
    std::vector<double> mutPos_;
    std::vector<std::string> mutAlleles_;
    std::vector<uint16_t> mutAltIndex_;
    bool anyMiss_ = false;
    double missPos_ = -1.0;
    uint32_t missNode_ = kInvalidNode;
    // Scratch reused across mutations (sized to the node / sample counts).
    std::vector<uint32_t> cnt_, stamp_, sampleMark_, ownerChild_;
    uint32_t epoch_ = 0;
};

} // namespace grgl_amd
