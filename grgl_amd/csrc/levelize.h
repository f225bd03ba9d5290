// Level-bucketed gather plans for the two sweeps of GRG matrix multiplication.
//
// Each direction gets its own renumbering of the nodes ("rows"): rows are grouped by
// topological level, so one level is a contiguous range of rows, and within a level the
// rows are sorted by gather-list length, longest first (equal lengths keep node order).
// Every row j has a gather list src[off[j] .. off[j+1]) of row indices whose values are
// summed, in list order, to give row j.  The list order is the reference CPU path's
// summation order (include/grgl/grg.h:1405-1474), so summing a list front to back
// reproduces the reference bit for bit:
//   UP   (samples -> mutations): row of node n sums the rows of n's children, ascending
//        child ID (matmulSumChildrenUp, grg.h:1324-1334).  Samples are level 0 and are
//        seeded from the input, never summed (grg.h:1471 starts at numSamples()).
//   DOWN (mutations -> samples): row of node n is its seeds followed by its parents in
//        descending parent ID (the push order of grg.h:1428-1431).  Seeds are the input
//        rows of its mutations (row numNodes + mutId) and of the mutations whose
//        missingness node it is (row numNodes + numMutations + mutId), interleaved in the
//        reference's (node, mutId) tuple order (grg.h:1406-1427).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "host_graph.h"

namespace grgl_amd {

struct DirPlan {
    uint32_t numNodes = 0;
    uint32_t numLevels = 0;
    uint32_t firstPulledLevel = 0;       // UP: 1 (level 0 is seeded), DOWN: 0
    std::vector<uint32_t> levelStart;    // numLevels + 1 row offsets
    std::vector<uint32_t> order;         // row -> node ID
    std::vector<uint32_t> rank;          // node ID -> row
    std::vector<uint32_t> off;           // numNodes + 1 gather-list offsets, modulo 2^32 (a
                                         // list's length is off[j + 1] - off[j] in u32)
    std::vector<uint32_t> src;           // gather-list entries (row indices)
    uint32_t maxFanIn = 0;
    // Readout: UP maps each mutation to its node's row (or invalid) and its missingness
    // node's row; DOWN maps each sample to its row.
    std::vector<uint32_t> outRow;
    std::vector<uint32_t> missRow;       // UP only, empty without missing data
    std::vector<uint32_t> coalByRow;     // per row: individual coalescence count (samples 0)
};

// Throws std::runtime_error if the graph is not an ordered GRG, one gather list exceeds 2^31
// entries, or (DOWN) nodes plus mutation seed rows exceed 2^32.
DirPlan buildUpPlan(const HostGrg& g);
DirPlan buildDownPlan(const HostGrg& g);

// Number of rows at the front of level `lvl` whose gather lists are longer than `heavy`.
uint32_t heavyPrefix(const DirPlan& p, uint32_t lvl, uint32_t heavy);

// Both plans of a graph and what the device graph needs besides them.
struct LevelizedGraph {
    uint32_t N = 0, V = 0, M = 0, P = 1;
    uint64_t E = 0;
    bool hasMiss = false, hasCoal = false;
    DirPlan up, down;
};

// Builds both plans (the two directions on two threads).  Throws std::runtime_error.
LevelizedGraph levelize(const HostGrg& g);

// Levelized-graph files: a header (magic, format version, graph sizes), both plans' arrays,
// and a 64-bit checksum of everything before it.  readLevelizedFile checks the checksum and
// every invariant the device kernels rely on (in-range and level-ordered gather entries,
// list lengths non-increasing within a level, permutations, read-out maps), so a file that
// loads cannot make a traversal read or write out of bounds.  Both throw std::runtime_error.
void writeLevelizedFile(const LevelizedGraph& lg, const std::string& path);
LevelizedGraph readLevelizedFile(const std::string& path);
void validateLevelized(const LevelizedGraph& lg);

} // namespace grgl_amd
