// Host-side immutable GRG in plain arrays: the content of the reference's CSRGRG
// (include/grgl/grg.h:1232-1300) without the varint encoding.  Nodes 0..numSamples-1 are
// the samples; every child ID is smaller than its parent's; children sorted ascending.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace grgl_amd {

constexpr uint32_t kInvalidNode = 0x7fffffffU;    // include/grgl/grgnode.h:58
constexpr uint32_t kCoalNotSet = 0xffffffffU;     // include/grgl/node_data.h:31

struct HostGrg {
    uint32_t numSamples = 0;
    uint32_t ploidy = 2;
    std::vector<uint64_t> downOff;   // numNodes + 1
    std::vector<uint32_t> downChild;
    std::vector<uint32_t> mutNode;   // per MutationId
    std::vector<uint32_t> mutMiss;   // per MutationId, empty when there is no missing data
    std::vector<double> mutPos;      // per MutationId (base-pair position)
    std::vector<uint32_t> coal;      // per node, empty when not available
    // Mutation metadata carried through .grg files (include/grgl/mutation.h:248-261):
    // allele storage string (REF then ALT, no delimiter), ALT start index, time field.
    std::vector<std::string> mutAlleles; // empty: "A" -> "G" for every mutation
    std::vector<uint16_t> mutAltIndex;
    std::vector<uint32_t> mutTime;
    std::vector<std::string> populations;
    std::vector<uint16_t> popIds;        // per sample, empty without populations
    std::vector<std::string> individualIds;
    uint64_t rangeStart = 0, rangeEnd = 0;
    bool phased = true;
    // Mutation IDs are in (position, REF, ALT) order as far as the graph's owner knows: files
    // read and immutable graphs built from arrays are; graphs grown by map_mutations are not,
    // and are written with their mutations sorted.
    bool mutationsOrdered = true;
    bool hasCoals() const { return !coal.empty(); }
    uint32_t numNodes() const { return downOff.empty() ? 0 : (uint32_t)(downOff.size() - 1); }
    uint64_t numEdges() const { return downOff.empty() ? 0 : downOff.back(); }
    uint32_t numMutations() const { return (uint32_t)mutNode.size(); }
};

// Throws std::runtime_error describing the first violated invariant.
void validateHostGrg(const HostGrg& g);

struct SynthParams {
    uint32_t numSamples = 1000;  // haplotypes
    uint32_t numSites = 10000;   // biallelic sites (= mutations)
    uint32_t numTrees = 1;       // independent coalescent trees, one per genome segment
    uint32_t ploidy = 2;
    uint64_t seed = 20261203ULL;
    uint32_t threads = 0;        // 0 = hardware concurrency
    double genomeLength = 1.0e8; // positions are spread over [0, genomeLength)
};

// Seeded Kingman coalescent trees (uniform pair merges, Exp(C(k,2)) inter-merge times),
// mutations on branches with probability proportional to branch length, each tree turned
// into GRG nodes with the reference's simplify-on-write rule (src/serialize.cpp:436-463:
// a single-parent node is dropped unless it carries a mutation; its children move up),
// nodes renumbered children-first.  Deterministic for a given SynthParams.
HostGrg synthCoalescentGrg(const SynthParams& p);

struct GraphStats {
    uint64_t numEdges = 0;
    uint32_t upLevels = 0;
    uint32_t maxFanIn = 0;
    std::vector<uint64_t> fanInHistogram; // bucket b counts nodes with fan-in in [2^b, 2^(b+1)), b=0 holds 0..1
};
GraphStats computeGraphStats(const HostGrg& g);

// The samples below each mutation's node (its carriers), sorted, as CSR: off has one entry per
// mutation plus one.  Samples reached by several paths are listed once.  The samples array
// comes in consecutive parts (one per worker thread), so a caller can move them into its own
// buffer one at a time.
struct CarrierParts {
    std::vector<uint64_t> off;
    std::vector<std::vector<uint32_t>> parts;
};
CarrierParts mutationCarriers(const HostGrg& g, uint32_t threads = 0);

// .grg files (format v5.x of the reference, src/serialize.cpp:518-669 and :697-890).
// Reading accepts versions 5.0-5.3 (u32 node index up to 5.1, u64 after); writing produces
// version 5.3, with the reference's simplify-on-write when allowSimplify, and mutations
// sorted by (position, REF, ALT) when the graph's are not marked ordered.
HostGrg readGrgFile(const std::string& path);
void writeGrgFile(const HostGrg& g, const std::string& path, bool allowSimplify);
// The graph as the reference's writer simplifies it (src/serialize.cpp:387-467).
HostGrg simplifyGrg(const HostGrg& g);

} // namespace grgl_amd
