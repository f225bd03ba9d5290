// The reference's graph constructor (fastGRGFromSamples, src/fast_build.cpp:794-986): GRG
// shape from haplotype similarity, mutations mapped onto it.  Input is the variant table as
// carrier lists (the reference reads VCF/IGD/BGEN through a MutationIterator; file readers are
// not part of this library), rows in genome order.
//
//   1. Rows are cut into trees.  Automatic: a tree ends once the unique haplotype segments of
//      its windows add up to a target that grows with log(samples) (getOptimalHapSegSum,
//      :736-752); with tree_count > 0, trees take equal numbers of rows.  Rows carried by fewer
//      samples than the no-tree threshold are mapped directly (one node, an edge per carrier).
//   2. Per tree, rows fall into windows of 254 variants; per window every sample's bits form a
//      haplotype segment, enumerated by uniqueness (createWindow, :136-182).
//   3. Nodes cluster bottom-up (buildTree, :430-706): each node finds its nearest neighbour by
//      Hamming distance summed over windows, in a BK-tree with a per-query comparison budget,
//      and the pair gives an intersection node (shared segments) plus nodes for what each side
//      holds alone; intersections join the index for the next pass.
//   4. Tree mapping (:656-706): each mutation goes onto the roots whose segments carry it
//      (a new node over them when there are several).
//   5. Trees are merged (grg_merge.cpp, node mapping by child set): a node whose mapped
//      children equal an existing node's is that node.
// Diploid graphs get coalescence counts (individuals whose two haplotypes arrive through
// different children).  Missing-data rows (allele ".", listing the haplotypes with no call)
// come first among their site's rows; they are mapped like the tree rows, and their node is
// the missingness node of the site's mutations.
#pragma once

#include <cstdint>
#include <vector>

#include "construct.h"

namespace grgl_amd {

struct FastBuildParams {
    uint32_t treeCount = 0;          // 0: automatic
    double noTreeBelow = 0.0;        // < 1: frequency, else count: rows below are mapped directly
    double rebuildProportion = 0.10; // rebuild the BK-tree when this share of its nodes is deleted
    bool treeMap = true;             // map tree rows onto the roots (else they stay unmapped)
    double hapSumMultiplier = 1.0;   // < 1 builds more, smaller trees (the reference's "faster" modes)
    uint32_t hapLength = 254;        // variants per window (at most 254)
};

struct FastBuildStats {
    uint32_t trees = 0;
    uint64_t windows = 0;
    uint64_t uniqueHapSegments = 0;  // summed over windows, before clustering
    uint64_t directMapped = 0;       // rows mapped directly
    uint64_t treeMapped = 0;         // rows mapped onto tree roots
    uint64_t unmapped = 0;           // rows with no carriers
    uint64_t comparisons = 0;        // BK-tree distance evaluations
    uint64_t queries = 0;
    uint64_t truncatedQueries = 0;   // queries that stopped on their comparison budget
    uint64_t rebuilds = 0;           // BK-tree rebuilds
    uint64_t combinedNodes = 0;      // nodes merged into an existing node of an earlier tree
};

// Trees are built on `threads` threads (0: one per hardware thread); the result does not
// depend on the thread count.  Mutation IDs follow the row order.
MutableGraph fastBuildFromSamples(uint32_t numSamples, uint32_t ploidy, const std::vector<MutationRecordIn>& rows,
                                  const std::vector<std::vector<uint32_t>>& carriers, const FastBuildParams& params,
                                  FastBuildStats& stats, uint32_t threads = 0, bool phased = true);
// The same with the carrier lists as CSR: row i's samples are samples[off[i] .. off[i+1]).
MutableGraph fastBuildFromSamples(uint32_t numSamples, uint32_t ploidy, const std::vector<MutationRecordIn>& rows,
                                  const uint64_t* off, const uint32_t* samples, const FastBuildParams& params,
                                  FastBuildStats& stats, uint32_t threads = 0, bool phased = true);

} // namespace grgl_amd
