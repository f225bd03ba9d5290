#include "grouping.h"

#include <algorithm>
#include <functional>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>

namespace grgl_amd {

namespace {

// Union-find over the examined rows.  Per set: value slots `a` (rows, 64 partials per heavy
// row, one per slot-list entry that reads a row outside the grouped levels: an upper bound on
// the imports), 32-bit words `b` (a list offset per row, a word per heavy row, one per
// slot-list entry) and gather entries; a block's LDS for class c is then at most
// (4 << c) * a + 4 * b plus its segment table.
struct Groups {
    std::vector<uint32_t> parent;
    std::vector<uint64_t> a, b, entries;
    explicit Groups(size_t n) : parent(n), a(n, 0), b(n, 0), entries(n, 0) {
        std::iota(parent.begin(), parent.end(), 0U);
    }
    uint32_t find(uint32_t x) {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    }
    uint32_t unite(uint32_t x, uint32_t y) {
        x = find(x);
        y = find(y);
        if (x == y) return x;
        if (a[x] < a[y]) std::swap(x, y);
        parent[y] = x;
        a[x] += a[y];
        b[x] += b[y];
        entries[x] += entries[y];
        return x;
    }
    uint64_t bytes(uint32_t root, uint32_t slotBytes) const { return a[root] * slotBytes + 4 * b[root]; }
};

// LDS every block needs besides its groups' share: the segment table (one segment per level
// at most), the closing list offset and alignment of the value slots.
uint64_t blockReserve(uint32_t levels) { return 4 * (1 + 3 * (uint64_t)levels) + 16; }

// The examined part of a plan: rows of levels [lowest, regionEnd), their lists, and the
// entries between them bucketed by the level of the row they read.
struct Examined {
    uint32_t lowest = 0, regionEnd = 0, rowBegin = 0, rowEnd = 0;
    std::vector<uint64_t> start;   // rows + 1 absolute list starts in p.src
    std::vector<uint32_t> level;   // per row
    std::vector<uint64_t> bucket;  // levels + 1 offsets into the edge arrays
    std::vector<uint32_t> edgeRow, edgeSrc, edgeIdx; // reader row, read row (from rowBegin), list index
};

uint64_t listLength(const DirPlan& p, uint32_t r) { return (uint32_t)(p.off[r + 1] - p.off[r]); }

Examined examine(const DirPlan& p, uint32_t regionEnd, const GroupLimits& lim) {
    Examined x;
    x.regionEnd = regionEnd;
    uint32_t lowest = regionEnd;
    uint64_t scanned = 0;
    while (lowest > p.firstPulledLevel) {
        uint64_t e = 0;
        for (uint32_t r = p.levelStart[lowest - 1]; r < p.levelStart[lowest]; r++) e += listLength(p, r);
        if (scanned + e > lim.maxScanEntries) break;
        scanned += e;
        lowest--;
    }
    x.lowest = lowest;
    x.rowBegin = p.levelStart[lowest];
    x.rowEnd = p.levelStart[regionEnd];
    const uint32_t n = x.rowEnd - x.rowBegin;
    uint64_t s0 = 0;
    for (uint32_t r = 0; r < x.rowBegin; r++) s0 += listLength(p, r);
    x.start.resize((size_t)n + 1);
    x.start[0] = s0;
    for (uint32_t i = 0; i < n; i++) x.start[i + 1] = x.start[i] + listLength(p, x.rowBegin + i);
    x.level.resize(n);
    for (uint32_t l = lowest; l < regionEnd; l++)
        for (uint32_t r = p.levelStart[l]; r < p.levelStart[l + 1]; r++) x.level[r - x.rowBegin] = l;
    const uint32_t nl = regionEnd - lowest;
    x.bucket.assign((size_t)nl + 1, 0);
    for (uint32_t i = 0; i < n; i++)
        for (uint64_t k = x.start[i]; k < x.start[i + 1]; k++) {
            const uint32_t s = p.src[k];
            if (s >= x.rowBegin && s < x.rowEnd) x.bucket[x.level[s - x.rowBegin] - lowest + 1]++;
        }
    for (uint32_t l = 0; l < nl; l++) x.bucket[l + 1] += x.bucket[l];
    x.edgeRow.resize(x.bucket[nl]);
    x.edgeSrc.resize(x.bucket[nl]);
    x.edgeIdx.resize(x.bucket[nl]);
    std::vector<uint64_t> cur(x.bucket.begin(), x.bucket.end() - 1);
    for (uint32_t i = 0; i < n; i++)
        for (uint64_t k = x.start[i]; k < x.start[i + 1]; k++) {
            const uint32_t s = p.src[k];
            if (s >= x.rowBegin && s < x.rowEnd) {
                const uint64_t o = cur[x.level[s - x.rowBegin] - lowest]++;
                x.edgeRow[o] = i;
                x.edgeSrc[o] = s - x.rowBegin;
                x.edgeIdx[o] = (uint32_t)(k - x.start[i]);
            }
        }
    return x;
}

// A row joins the examined levels with no in-group entry yet (its sources are all in lower
// levels): its slots, words and entries.
void initGroups(const DirPlan& p, const Examined& x, uint32_t heavyFanin, Groups& g, std::vector<uint32_t>& pre) {
    const uint32_t n = x.rowEnd - x.rowBegin;
    pre.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t len = listLength(p, x.rowBegin + i);
        const bool heavy = len > heavyFanin;
        g.a[i] = heavy ? 1 + kGroupPartials : 1;
        g.b[i] = heavy ? 2 : 1;
        g.entries[i] = len;
        pre[i] = (uint32_t)len;
    }
}

// Adds the entries of level L's bucket: unions, and each reader row's slot list growing to
// start at its first in-group entry.  `touch` sees every changed root.
template <typename Touch>
void addLevel(const Examined& x, uint32_t L, Groups& g, std::vector<uint32_t>& pre, Touch touch,
              uint64_t* totalA = nullptr, uint64_t* totalB = nullptr) {
    for (uint64_t k = x.bucket[L - x.lowest]; k < x.bucket[L - x.lowest + 1]; k++) {
        const uint32_t r = x.edgeRow[k];
        const uint32_t root = g.unite(r, x.edgeSrc[k]);
        const uint32_t idx = x.edgeIdx[k];
        // The entry becomes an in-group slot (no import); the slot list grows to start at it.
        const uint64_t d = idx < pre[r] ? pre[r] - idx : 0;
        if (d) pre[r] = idx;
        g.a[root] += d;
        g.a[root] -= 1;
        g.b[root] += d;
        if (totalA) *totalA += d - 1;
        if (totalB) *totalB += d;
        touch(root);
    }
}

// Groups of levels [l0, regionEnd) (the union-find holds exactly their entries) packed into
// blocks, and the block programs built from them.
void buildPlan(const DirPlan& p, const Examined& x, uint32_t l0, uint32_t heavyFanin, uint32_t slotBytes,
               const GroupLimits& lim, Groups& g, GroupPlan& gp) {
    const uint32_t Rb = p.levelStart[l0], R1 = x.rowEnd, m = R1 - Rb, base = Rb - x.rowBegin;
    const uint32_t nl = x.regionEnd - l0;
    const uint64_t capBytes = lim.ldsBytes - blockReserve(nl);
    // Dense group ids in order of first row.
    std::vector<uint32_t> groupOf(m), idOfRoot(x.rowEnd - x.rowBegin, UINT32_MAX);
    std::vector<uint64_t> gBytes, gCost;
    for (uint32_t i = 0; i < m; i++) {
        const uint32_t root = g.find(base + i);
        if (idOfRoot[root] == UINT32_MAX) {
            idOfRoot[root] = (uint32_t)gBytes.size();
            gBytes.push_back(g.bytes(root, slotBytes));
            gCost.push_back(g.a[root] + g.entries[root]);
        }
        groupOf[i] = idOfRoot[root];
    }
    const uint32_t nGroups = (uint32_t)gBytes.size();
    // Longest-processing-time packing: largest group first into the least-loaded block that
    // has LDS left (a new block when the least-loaded one has not).
    std::vector<uint32_t> byCost(nGroups);
    std::iota(byCost.begin(), byCost.end(), 0U);
    std::stable_sort(byCost.begin(), byCost.end(), [&](uint32_t a, uint32_t b) { return gCost[a] > gCost[b]; });
    const uint32_t want = std::min<uint32_t>(nGroups, std::max<uint32_t>(1, lim.maxBlocks));
    using Load = std::pair<uint64_t, uint32_t>;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    std::vector<uint64_t> blockBytes(want, 0);
    for (uint32_t b = 0; b < want; b++) heap.push({0, b});
    std::vector<uint32_t> blockOfGroup(nGroups);
    for (uint32_t gi : byCost) {
        Load top = heap.top();
        heap.pop();
        uint32_t b = top.second;
        if (blockBytes[b] + gBytes[gi] > capBytes) {
            heap.push(top);
            b = (uint32_t)blockBytes.size();
            blockBytes.push_back(0);
            top = {0, b};
        }
        blockOfGroup[gi] = b;
        blockBytes[b] += gBytes[gi];
        heap.push({top.first + gCost[gi], b});
    }
    // Every block holds a group: the first `want` groups go to the empty blocks one each.
    const uint32_t B = (uint32_t)blockBytes.size();
    gp.blocks = B;
    gp.levelBegin = l0;
    gp.levelEnd = x.regionEnd;
    gp.heavyFanin = heavyFanin;
    gp.slotBytes = slotBytes;
    // Rows to blocks, in plan order (level by level, longest list first within a level).
    gp.blockRow.assign((size_t)B + 1, 0);
    for (uint32_t i = 0; i < m; i++) gp.blockRow[blockOfGroup[groupOf[i]] + 1]++;
    for (uint32_t b = 0; b < B; b++) gp.blockRow[b + 1] += gp.blockRow[b];
    gp.row.resize(m);
    std::vector<uint32_t> slotOf(m);
    {
        std::vector<uint32_t> cur(gp.blockRow.begin(), gp.blockRow.end() - 1);
        for (uint32_t i = 0; i < m; i++) {
            const uint32_t b = blockOfGroup[groupOf[i]];
            const uint32_t pos = cur[b]++;
            gp.row[pos] = Rb + i;
            slotOf[i] = pos - gp.blockRow[b];
        }
    }
    // Segments (one level of one block) and heavy rows.
    gp.blockSeg.assign((size_t)B + 1, 0);
    gp.blockHeavy.assign((size_t)B + 1, 0);
    gp.segRow.clear();
    gp.segHeavy.clear();
    gp.segHeavyOrd.clear();
    gp.heavyRow.clear();
    for (uint32_t b = 0; b < B; b++) {
        gp.blockSeg[b] = (uint32_t)gp.segHeavy.size();
        gp.blockHeavy[b] = (uint32_t)gp.heavyRow.size();
        uint32_t curLevel = UINT32_MAX, ord = 0;
        for (uint32_t pos = gp.blockRow[b]; pos < gp.blockRow[b + 1]; pos++) {
            const uint32_t r = gp.row[pos];
            const uint32_t lvl = x.level[r - x.rowBegin];
            if (lvl != curLevel) {
                curLevel = lvl;
                gp.segRow.push_back(pos);
                gp.segHeavy.push_back(0);
                gp.segHeavyOrd.push_back(ord);
            }
            if (listLength(p, r) > heavyFanin) {
                if (gp.segRow.back() + gp.segHeavy.back() != pos)
                    throw std::logic_error("grouping: heavy rows not at the front of a level");
                gp.segHeavy.back()++;
                gp.heavyRow.push_back(pos);
                ord++;
            }
        }
    }
    gp.blockSeg[B] = (uint32_t)gp.segHeavy.size();
    gp.blockHeavy[B] = (uint32_t)gp.heavyRow.size();
    gp.segRow.push_back(m);
    // Lists: plan rows outside the grouped levels, tagged rows inside; slot lists from each
    // row's first in-group entry on, HBM rows there read through the block's import slots.
    gp.off.assign((size_t)m + 1, 0);
    gp.pre.assign(m, 0);
    gp.remOff.assign((size_t)m + 1, 0);
    uint64_t total = 0;
    for (uint32_t pos = 0; pos < m; pos++) total += listLength(p, gp.row[pos]);
    if (total >= (1ULL << 32)) throw std::logic_error("grouping: too many entries");
    gp.src.resize(total);
    gp.rem.clear();
    gp.importRow.clear();
    gp.blockImport.assign((size_t)B + 1, 0);
    gp.ldsBytes = 0;
    uint64_t o = 0;
    std::unordered_map<uint32_t, uint32_t> importOf;
    for (uint32_t b = 0; b < B; b++) {
        gp.blockImport[b] = (uint32_t)gp.importRow.size();
        importOf.clear();
        const uint32_t nRows = gp.blockRow[b + 1] - gp.blockRow[b];
        const uint32_t nHeavy = gp.blockHeavy[b + 1] - gp.blockHeavy[b];
        const uint32_t importBase = nRows + kGroupPartials * nHeavy;
        const uint64_t remBegin = gp.rem.size();
        for (uint32_t pos = gp.blockRow[b]; pos < gp.blockRow[b + 1]; pos++) {
            const uint32_t i = gp.row[pos] - x.rowBegin;
            gp.off[pos] = (uint32_t)o;
            gp.remOff[pos] = (uint32_t)gp.rem.size();
            const uint64_t len = x.start[i + 1] - x.start[i];
            uint32_t pre = (uint32_t)len;
            for (uint64_t k = 0; k < len; k++) {
                const uint32_t s = p.src[x.start[i] + k];
                const bool inGroup = s >= Rb && s < R1;
                if (inGroup && pre == len) pre = (uint32_t)k;
                if (inGroup) {
                    gp.src[o++] = kGroupSlotTag | slotOf[s - Rb];
                } else {
                    if (s & kGroupSlotTag) throw std::logic_error("grouping: row index needs the slot bit");
                    gp.src[o++] = s;
                }
                if (k < pre) continue;
                if (inGroup) {
                    gp.rem.push_back(slotOf[s - Rb]);
                } else {
                    auto it = importOf.find(s);
                    if (it == importOf.end()) {
                        it = importOf.emplace(s, (uint32_t)gp.importRow.size() - gp.blockImport[b]).first;
                        gp.importRow.push_back(s);
                    }
                    gp.rem.push_back(importBase + it->second);
                }
            }
            gp.pre[pos] = pre;
        }
        const uint64_t lds = groupLdsBytes(slotBytes, nRows, nHeavy, gp.importRow.size() - gp.blockImport[b],
                                           gp.rem.size() - remBegin, gp.blockSeg[b + 1] - gp.blockSeg[b]);
        gp.ldsBytes = std::max<uint32_t>(gp.ldsBytes, (uint32_t)lds);
    }
    gp.blockImport[B] = (uint32_t)gp.importRow.size();
    gp.off[m] = (uint32_t)o;
    gp.remOff[m] = (uint32_t)gp.rem.size();
    if (gp.ldsBytes > lim.ldsBytes) throw std::logic_error("grouping: a block exceeds its LDS");
}

} // namespace

std::vector<GroupPlan> buildGroupPlans(const DirPlan& p, uint32_t regionEnd, uint32_t heavyFanin,
                                       const GroupLimits& lim) {
    std::vector<GroupPlan> out(kGroupClasses);
    for (uint32_t c = 0; c < kGroupClasses; c++) {
        out[c].slotBytes = groupClassBytes(c);
        out[c].heavyFanin = heavyFanin;
    }
    regionEnd = std::min(regionEnd, p.numLevels);
    if (regionEnd < p.firstPulledLevel + 2) return out;
    const Examined x = examine(p, regionEnd, lim);
    const uint32_t nl = regionEnd - x.lowest;
    if (nl < 2) return out;
    const uint32_t n = x.rowEnd - x.rowBegin;
    // Pass 1: the largest group of levels [L, regionEnd) for each L, in LDS bytes per class
    // and in gather entries.
    // Also the total of all groups' slots and words, whose LDS must fit the resident blocks.
    std::vector<std::vector<uint64_t>> maxBytes(kGroupClasses, std::vector<uint64_t>(nl));
    std::vector<uint64_t> maxEntries(nl), totalA(nl), totalB(nl);
    {
        Groups g(n);
        std::vector<uint32_t> pre;
        initGroups(p, x, heavyFanin, g, pre);
        std::vector<uint64_t> mb(kGroupClasses, 0);
        uint64_t me = 0, ta = 0, tb = 0;
        auto touch = [&](uint32_t root) {
            for (uint32_t c = 0; c < kGroupClasses; c++) mb[c] = std::max(mb[c], g.bytes(root, groupClassBytes(c)));
            me = std::max(me, g.entries[root]);
        };
        for (uint32_t L = regionEnd; L-- > x.lowest;) {
            for (uint32_t r = p.levelStart[L]; r < p.levelStart[L + 1]; r++) {
                touch(r - x.rowBegin);
                ta += g.a[r - x.rowBegin];
                tb += g.b[r - x.rowBegin];
            }
            addLevel(x, L, g, pre, touch, &ta, &tb);
            for (uint32_t c = 0; c < kGroupClasses; c++) maxBytes[c][L - x.lowest] = mb[c];
            maxEntries[L - x.lowest] = me;
            totalA[L - x.lowest] = ta;
            totalB[L - x.lowest] = tb;
        }
    }
    // First grouped level per class: the lowest L whose groups all fit.
    std::vector<uint32_t> first(kGroupClasses, regionEnd);
    uint32_t lowestWanted = regionEnd;
    for (uint32_t c = 0; c < kGroupClasses; c++) {
        if (!((lim.classMask >> c) & 1U)) continue;
        uint32_t l0 = regionEnd;
        while (l0 > x.lowest) {
            const uint32_t L = l0 - 1;
            const uint64_t cap = lim.ldsBytes - blockReserve(regionEnd - L);
            const uint64_t total = totalA[L - x.lowest] * groupClassBytes(c) + 4 * totalB[L - x.lowest];
            // Packing leaves blocks partly empty: the groups' total must fit 85% of the resident
            // blocks' LDS.
            if (maxBytes[c][L - x.lowest] + blockReserve(regionEnd - L) > lim.ldsBytes ||
                maxEntries[L - x.lowest] > lim.maxEntries || total * 100 > cap * lim.maxBlocks * 85 ||
                p.levelStart[L + 1] - p.levelStart[L] > lim.maxLevelRows)
                break;
            l0 = L;
        }
        if (regionEnd - l0 >= 2) {
            first[c] = l0;
            lowestWanted = std::min(lowestWanted, l0);
        }
    }
    if (lowestWanted == regionEnd) return out;
    // Pass 2: the same unions, each class's groups taken at its first level.
    Groups g(n);
    std::vector<uint32_t> pre;
    initGroups(p, x, heavyFanin, g, pre);
    for (uint32_t L = regionEnd; L-- > lowestWanted;) {
        addLevel(x, L, g, pre, [](uint32_t) {});
        for (uint32_t c = 0; c < kGroupClasses; c++)
            if (first[c] == L) buildPlan(p, x, L, heavyFanin, groupClassBytes(c), lim, g, out[c]);
    }
    return out;
}

void validateGroupPlan(const DirPlan& p, const GroupPlan& gp, const GroupLimits& lim) {
    auto bad = [](const std::string& why) { throw std::runtime_error("invalid group plan: " + why); };
    if (gp.empty()) return;
    if (gp.levelEnd > p.numLevels || gp.levelBegin < p.firstPulledLevel || gp.levelEnd < gp.levelBegin + 2)
        bad("levels");
    for (uint32_t l = gp.levelBegin; l < gp.levelEnd; l++)
        if (p.levelStart[l + 1] - p.levelStart[l] > lim.maxLevelRows) bad("a level over the row limit");
    const uint32_t Rb = p.levelStart[gp.levelBegin], R1 = p.levelStart[gp.levelEnd], m = R1 - Rb;
    const uint32_t B = gp.blocks;
    if (gp.row.size() != m || gp.off.size() != (size_t)m + 1 || gp.pre.size() != m ||
        gp.remOff.size() != (size_t)m + 1 || gp.blockRow.size() != (size_t)B + 1 ||
        gp.blockSeg.size() != (size_t)B + 1 || gp.blockHeavy.size() != (size_t)B + 1 ||
        gp.blockImport.size() != (size_t)B + 1 || gp.segRow.size() != gp.segHeavy.size() + 1 ||
        gp.segHeavyOrd.size() != gp.segHeavy.size() || gp.src.size() != gp.off[m] || gp.off[0] != 0 ||
        gp.rem.size() != gp.remOff[m] || gp.remOff[0] != 0)
        bad("sizes");
    if (gp.blockRow[0] != 0 || gp.blockRow[B] != m || gp.blockSeg[0] != 0 || gp.blockSeg[B] != gp.segHeavy.size() ||
        gp.blockHeavy[0] != 0 || gp.blockHeavy[B] != gp.heavyRow.size() || gp.blockImport[0] != 0 ||
        gp.blockImport[B] != gp.importRow.size())
        bad("block bounds");
    std::vector<uint32_t> seen(m, 0), blockOf(m), slotOf(m), levelOfRow(m);
    for (uint32_t l = gp.levelBegin; l < gp.levelEnd; l++)
        for (uint32_t r = p.levelStart[l]; r < p.levelStart[l + 1]; r++) levelOfRow[r - Rb] = l;
    uint32_t maxLds = 0;
    for (uint32_t b = 0; b < B; b++) {
        if (gp.blockRow[b + 1] < gp.blockRow[b] || gp.blockSeg[b + 1] < gp.blockSeg[b] ||
            gp.blockHeavy[b + 1] < gp.blockHeavy[b] || gp.blockImport[b + 1] < gp.blockImport[b])
            bad("block bounds");
        uint32_t heavySeen = 0, prevLevel = 0;
        for (uint32_t s = gp.blockSeg[b]; s < gp.blockSeg[b + 1]; s++) {
            const uint32_t s0 = gp.segRow[s], s1 = gp.segRow[s + 1];
            if (s0 >= s1 || (s == gp.blockSeg[b] && s0 != gp.blockRow[b]) ||
                (s + 1 == gp.blockSeg[b + 1] && s1 != gp.blockRow[b + 1]))
                bad("segment bounds");
            if (gp.segHeavy[s] > s1 - s0 || gp.segHeavyOrd[s] != heavySeen) bad("segment heavy rows");
            for (uint32_t pos = s0; pos < s1; pos++) {
                const uint32_t r = gp.row[pos];
                if (r < Rb || r >= R1 || seen[r - Rb]++) bad("rows are not the grouped levels once each");
                blockOf[r - Rb] = b;
                slotOf[r - Rb] = pos - gp.blockRow[b];
                const uint32_t lvl = levelOfRow[r - Rb];
                if ((pos == s0 && s != gp.blockSeg[b] && lvl <= prevLevel) ||
                    (pos > s0 && lvl != levelOfRow[gp.row[pos - 1] - Rb]))
                    bad("segments not level-ordered");
                prevLevel = lvl;
                const bool heavy = (uint32_t)(p.off[r + 1] - p.off[r]) > gp.heavyFanin;
                if (heavy != (pos - s0 < gp.segHeavy[s])) bad("heavy rows");
                if (heavy) {
                    if (gp.blockHeavy[b] + heavySeen >= gp.blockHeavy[b + 1] ||
                        gp.heavyRow[gp.blockHeavy[b] + heavySeen] != pos)
                        bad("heavy row list");
                    heavySeen++;
                }
            }
        }
        if (gp.blockHeavy[b] + heavySeen != gp.blockHeavy[b + 1]) bad("heavy row list");
        const uint32_t nRows = gp.blockRow[b + 1] - gp.blockRow[b];
        const uint64_t lds = groupLdsBytes(gp.slotBytes, nRows, heavySeen, gp.blockImport[b + 1] - gp.blockImport[b],
                                           gp.remOff[gp.blockRow[b + 1]] - gp.remOff[gp.blockRow[b]],
                                           gp.blockSeg[b + 1] - gp.blockSeg[b]);
        if (lds > lim.ldsBytes) bad("LDS");
        maxLds = std::max<uint32_t>(maxLds, (uint32_t)lds);
    }
    if (maxLds != gp.ldsBytes) bad("LDS");
    // Lists: the plan's, with the grouped levels' rows tagged; slot lists read the same rows.
    std::vector<uint64_t> start((size_t)p.numNodes + 1, 0);
    for (uint32_t r = 0; r < p.numNodes; r++) start[r + 1] = start[r] + (uint32_t)(p.off[r + 1] - p.off[r]);
    for (uint32_t pos = 0; pos < m; pos++) {
        const uint32_t r = gp.row[pos], b = blockOf[r - Rb];
        const uint32_t nRows = gp.blockRow[b + 1] - gp.blockRow[b];
        const uint32_t nHeavy = gp.blockHeavy[b + 1] - gp.blockHeavy[b];
        const uint32_t nImport = gp.blockImport[b + 1] - gp.blockImport[b];
        const uint32_t importBase = nRows + kGroupPartials * nHeavy;
        const uint64_t len = start[r + 1] - start[r];
        if ((uint64_t)gp.off[pos + 1] - gp.off[pos] != len) bad("list length");
        uint32_t pre = (uint32_t)len;
        for (uint64_t k = 0; k < len; k++) {
            const uint32_t s = p.src[start[r] + k], e = gp.src[gp.off[pos] + k];
            const bool inGroup = s >= Rb && s < R1;
            if (inGroup) {
                if (pre == len) pre = (uint32_t)k;
                if (blockOf[s - Rb] != b) bad("a group reads another block's row");
                if (e != (kGroupSlotTag | slotOf[s - Rb])) bad("slot entry");
            } else if (e != s) {
                bad("row entry");
            }
        }
        if (gp.pre[pos] != pre) bad("list prefix");
        if ((uint64_t)gp.remOff[pos + 1] - gp.remOff[pos] != len - pre) bad("slot list length");
        for (uint64_t k = pre; k < len; k++) {
            const uint32_t s = p.src[start[r] + k], slot = gp.rem[gp.remOff[pos] + (k - pre)];
            if (s >= Rb && s < R1) {
                if (slot != slotOf[s - Rb]) bad("slot list entry");
            } else if (slot < importBase || slot >= importBase + nImport ||
                       gp.importRow[gp.blockImport[b] + (slot - importBase)] != s) {
                bad("import slot");
            }
        }
    }
}

} // namespace grgl_amd
