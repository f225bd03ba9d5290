// Grouped launches: the later levels of a traversal plan cut into connected groups of rows
// that one block traverses by itself, level after level with block barriers in between.
//
// A group is a connected component of the rows of levels [levelBegin, levelEnd) under the
// gather entries between them: no row of the group reads a row of those levels outside it.
// So once the levels before levelBegin (and the seed rows) are complete, each block can run
// its groups to the end without waiting for any other block.  Several groups share a block
// (balanced by work).  A block first loads everything its levels need into LDS in one
// parallel pass (phase 0): each row's leading entries that read rows in HBM are summed there
// and then, the rest of every list becomes a list of LDS slots (the block's own rows, and
// "import" slots holding the other HBM rows read after a row's first in-group entry), and the
// segment table is copied too.  The level-by-level part (phase 1) then reads only LDS.
//
// Every row keeps its gather list in the plan's order and its split into 64 partials when
// it is longer than the heavy threshold, so a grouped traversal computes every row with the
// same arithmetic as the level kernels: identical bits.
#pragma once

#include <cstdint>
#include <vector>

#include "levelize.h"

namespace grgl_amd {

// A list entry of `src` with this bit set reads a row of the block's groups (its value
// comes from the block's LDS; `rem` holds the slot); plan rows and seed rows stay below 2^31.
constexpr uint32_t kGroupSlotTag = 0x80000000U;
// Row classes of grouped launches: slots of 4, 8, 16, 32 or 64 bytes.
constexpr uint32_t kGroupClasses = 5;
inline uint32_t groupClassBytes(uint32_t c) { return 4U << c; }
constexpr uint32_t kGroupPartials = 64; // partial slots of a heavy row (the level kernels' split)

struct GroupLimits {
    uint32_t ldsBytes = 64U << 10;   // LDS of one block at most
    uint32_t maxEntries = 16384;     // gather entries of one group at most
    // Blocks of a grouped launch at most: as many as are resident at once, so that no block
    // waits for a compute unit (MI355X: 256 compute units of 160 KB LDS, two 64 KB blocks
    // each).  The grouped levels stop where their groups would need more LDS than that.
    uint32_t maxBlocks = 512;
    // Gather entries examined per direction at most: bounds the plans' host time and device
    // memory (each class's list copies, 8 bytes an entry) on graphs of billions of edges.
    uint64_t maxScanEntries = 1ULL << 24;
    // Grouped levels hold at most this many rows each: a lower bound leaves the larger levels
    // to the level kernels (which move wide rows faster) and groups only the small levels after
    // them.
    uint32_t maxLevelRows = 0xffffffffU;
    // Row classes built (bit c: class c).
    uint32_t classMask = (1U << kGroupClasses) - 1;
};

// LDS of one block, in bytes: value slots (rows, 64 partials per heavy row, imports) of
// `slotBytes` each, then 32-bit words: the rows' offsets into the slot lists (rows + 1), the
// slot lists, three words per segment and one per heavy row.
inline uint64_t groupLdsBytes(uint32_t slotBytes, uint64_t rows, uint64_t heavy, uint64_t imports, uint64_t remEntries,
                              uint64_t segments) {
    const uint64_t values = (rows + kGroupPartials * heavy + imports) * slotBytes;
    return (values + 15) / 16 * 16 + 4 * (rows + 1 + remEntries + 3 * segments + heavy);
}

struct GroupPlan {
    uint32_t slotBytes = 0;                  // 4 << class
    uint32_t heavyFanin = 0;                 // rows with longer lists hold 64 partial slots
    uint32_t levelBegin = 0, levelEnd = 0;   // grouped levels; none when equal
    uint32_t blocks = 0;
    uint32_t ldsBytes = 0;                   // LDS of the largest block
    // Blocks (blocks + 1 offsets each): first group row, segment, heavy row and import.
    std::vector<uint32_t> blockRow, blockSeg, blockHeavy, blockImport;
    // Segments (one level of one block): first group row (segments + 1), and the heavy rows at
    // its front with the block-relative ordinal of the first of them.
    std::vector<uint32_t> segRow, segHeavy, segHeavyOrd;
    // Group rows, block after block, each block's rows level by level in plan order (heavy
    // rows first within a level): the plan row computed; its list (off: rows + 1 bounds in src,
    // entries as in the plan, in-group ones tagged); `pre` leading entries read HBM rows
    // (summed in phase 0); the rest is the row's slot list rem[remOff[i] .. remOff[i + 1]).
    std::vector<uint32_t> row, off, pre, src, remOff, rem;
    std::vector<uint32_t> heavyRow;          // each block's heavy group rows, in segment order
    std::vector<uint32_t> importRow;         // each block's import slots: the HBM row each holds
    bool empty() const { return levelBegin >= levelEnd; }
};

// Where grouped levels end: UP, the last level; DOWN, before the last level, which holds the
// rows without children (every sample among them): a sample reads one row of each tree
// above it and would join all their groups into one.
inline uint32_t groupRegionEnd(const DirPlan& p, bool down) {
    return down && p.numLevels > 0 ? p.numLevels - 1 : p.numLevels;
}

// Grouped levels of `p` for each row class: the longest run of levels ending at `regionEnd`
// whose groups fit the limits at the class's slot size, when it spans at least two levels.
// Slot c of the result is class c.
std::vector<GroupPlan> buildGroupPlans(const DirPlan& p, uint32_t regionEnd, uint32_t heavyFanin,
                                       const GroupLimits& lim);

// Throws std::runtime_error unless `gp` is a grouping of levels [levelBegin, levelEnd) of
// `p`: every row of those levels exactly once, level-ordered segments, lists equal to the
// plan's with in-group entries as slots of the same block, slot lists and imports that
// read the same rows, heavy rows and LDS as stated, levels within lim.maxLevelRows.
void validateGroupPlan(const DirPlan& p, const GroupPlan& gp, const GroupLimits& lim);

} // namespace grgl_amd
