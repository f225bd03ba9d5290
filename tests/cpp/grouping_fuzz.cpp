// Grouped-level plans under random graphs and limits (run by tests/test_grouping_fuzz.py,
// built with the address and undefined-behaviour sanitizers): for every direction of a random
// ordered graph and random limits (LDS per block, entries per group, blocks, heavy threshold,
// entries examined, rows per level, classes built), every plan buildGroupPlans returns must
// pass validateGroupPlan, fit the
// LDS limit, and compute — executed on the CPU the way pull_group_kernel does it (leading HBM
// entries first, then each level's rows from LDS slots, heavy rows as 64 partials and a fixed
// tree) — the same float bits as the level plan executed the way the level kernels do it.
// Exit code 0 = all checks held.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "grouping.h"
#include "levelize.h"

using namespace grgl_amd;

static HostGrg randomGraph(std::mt19937_64& rng) {
    HostGrg g;
    const uint32_t N = 2 * (2 + (uint32_t)(rng() % 30));
    g.numSamples = N;
    g.ploidy = 2;
    g.downOff.assign(N + 1, 0);
    const bool forest = rng() % 2 == 0;
    const uint32_t internal = 1 + (uint32_t)(rng() % 200);
    for (uint32_t v = N; v < N + internal; v++) {
        std::vector<uint32_t> kids;
        if (forest && v > N + 3) {
            // tree-like: a few recent nodes and a few samples
            const uint32_t k = 1 + (uint32_t)(rng() % 3);
            for (uint32_t i = 0; i < k; i++) kids.push_back(N + (uint32_t)(rng() % (v - N)));
            const uint32_t s = (uint32_t)(rng() % (rng() % 5 == 0 ? 100 : 3));
            for (uint32_t i = 0; i < s; i++) kids.push_back((uint32_t)(rng() % N));
        } else {
            const uint32_t want = 1 + (uint32_t)(rng() % (rng() % 4 == 0 ? 150 : 5));
            for (uint32_t i = 0; i < want; i++) kids.push_back((uint32_t)(rng() % v));
        }
        std::sort(kids.begin(), kids.end());
        kids.erase(std::unique(kids.begin(), kids.end()), kids.end());
        g.downChild.insert(g.downChild.end(), kids.begin(), kids.end());
        g.downOff.push_back(g.downChild.size());
    }
    const uint32_t V = N + internal;
    const uint32_t M = (uint32_t)(rng() % 300);
    for (uint32_t m = 0; m < M; m++) g.mutNode.push_back(rng() % 9 == 0 ? kInvalidNode : (uint32_t)(rng() % V));
    if (rng() % 3 == 0)
        for (uint32_t m = 0; m < M; m++) g.mutMiss.push_back(rng() % 3 ? kInvalidNode : (uint32_t)(rng() % V));
    for (uint32_t m = 0; m < M; m++) g.mutPos.push_back((double)m);
    return g;
}

static const int K = 2;
using Row = std::vector<float>;

static Row zeroRow() { return Row(K, 0.0f); }
static void addInto(Row& a, const Row& b) {
    for (int k = 0; k < K; k++) a[k] = a[k] + b[k];
}

// One row as the level kernels compute it.
static Row sumRow(const std::vector<Row>& rows, const uint32_t* lst, uint64_t len, uint32_t heavy) {
    if (len <= heavy) {
        Row acc = zeroRow();
        for (uint64_t i = 0; i < len; i++) addInto(acc, rows[lst[i]]);
        return acc;
    }
    std::vector<Row> part(kGroupPartials, zeroRow());
    for (uint64_t i = 0; i < len; i++) addInto(part[i % kGroupPartials], rows[lst[i]]);
    for (uint32_t st = kGroupPartials / 2; st >= 1; st /= 2)
        for (uint32_t t = 0; t < st; t++) addInto(part[t], part[t + st]);
    return part[0];
}

static void runLevels(std::vector<Row>& rows, const DirPlan& p, uint32_t lo, uint32_t hi, uint32_t heavy,
                      const std::vector<uint64_t>& start) {
    for (uint32_t l = lo; l < hi; l++)
        for (uint32_t j = p.levelStart[l]; j < p.levelStart[l + 1]; j++)
            rows[j] = sumRow(rows, p.src.data() + start[j], start[j + 1] - start[j], heavy);
}

// pull_group_kernel on the CPU, block after block.
static void runGroups(std::vector<Row>& rows, const GroupPlan& gp) {
    for (uint32_t b = 0; b < gp.blocks; b++) {
        const uint32_t r0 = gp.blockRow[b], r1 = gp.blockRow[b + 1];
        const uint32_t h0 = gp.blockHeavy[b], h1 = gp.blockHeavy[b + 1];
        const uint32_t i0 = gp.blockImport[b], i1 = gp.blockImport[b + 1];
        const uint32_t n = r1 - r0, nh = h1 - h0;
        const uint32_t importBase = n + kGroupPartials * nh;
        std::vector<Row> slots(importBase + (i1 - i0), zeroRow());
        for (uint32_t t = 0; t < i1 - i0; t++) slots[importBase + t] = rows.at(gp.importRow[i0 + t]);
        for (uint32_t i = r0; i < r1; i++) {
            const uint32_t lb = gp.off[i], le = gp.off[i + 1];
            if (le - lb > gp.heavyFanin) continue;
            Row acc = zeroRow();
            for (uint32_t k = lb; k < lb + gp.pre[i]; k++) {
                if (gp.src[k] & kGroupSlotTag) throw std::runtime_error("prefix entry is a slot");
                addInto(acc, rows.at(gp.src[k]));
            }
            slots[i - r0] = acc;
        }
        for (uint32_t k = 0; k < nh; k++) {
            const uint32_t i = gp.heavyRow[h0 + k];
            const uint32_t lb = gp.off[i];
            for (uint32_t q = 0; q < kGroupPartials; q++) {
                Row acc = zeroRow();
                for (uint32_t t = lb + q; t < lb + gp.pre[i]; t += kGroupPartials) addInto(acc, rows.at(gp.src[t]));
                slots[n + kGroupPartials * k + q] = acc;
            }
        }
        for (uint32_t s = gp.blockSeg[b]; s < gp.blockSeg[b + 1]; s++) {
            const uint32_t g0 = gp.segRow[s], g1 = gp.segRow[s + 1], hc = gp.segHeavy[s], hord = gp.segHeavyOrd[s];
            for (uint32_t h = 0; h < hc; h++) {
                const uint32_t i = g0 + h, k = hord + h, p = gp.pre[i];
                std::vector<Row> part(kGroupPartials);
                for (uint32_t q = 0; q < kGroupPartials; q++) part[q] = slots.at(n + kGroupPartials * k + q);
                for (uint32_t j = gp.remOff[i]; j < gp.remOff[i + 1]; j++)
                    addInto(part[(p + (j - gp.remOff[i])) % kGroupPartials], slots.at(gp.rem[j]));
                for (uint32_t st = kGroupPartials / 2; st >= 1; st /= 2)
                    for (uint32_t t = 0; t < st; t++) addInto(part[t], part[t + st]);
                slots[i - r0] = part[0];
            }
            for (uint32_t i = g0 + hc; i < g1; i++) {
                Row acc = slots[i - r0];
                for (uint32_t j = gp.remOff[i]; j < gp.remOff[i + 1]; j++) addInto(acc, slots.at(gp.rem[j]));
                slots[i - r0] = acc;
            }
        }
        for (uint32_t i = r0; i < r1; i++) rows.at(gp.row[i]) = slots[i - r0];
    }
}

int main(int argc, char** argv) {
    const int graphs = argc > 1 ? atoi(argv[1]) : 100;
    std::mt19937_64 rng(20261204);
    uint64_t plansChecked = 0;
    for (int it = 0; it < graphs; it++) {
        const HostGrg g = randomGraph(rng);
        const LevelizedGraph lg = levelize(g);
        for (int dir = 0; dir < 2; dir++) {
            const bool down = dir == 1;
            const DirPlan& p = down ? lg.down : lg.up;
            std::vector<uint64_t> start(p.numNodes + 1, 0);
            for (uint32_t r = 0; r < p.numNodes; r++) start[r + 1] = start[r] + (uint32_t)(p.off[r + 1] - p.off[r]);
            const uint32_t numRows = p.numNodes + (down ? lg.M * (lg.hasMiss ? 2 : 1) : 0);
            std::vector<Row> seeded(numRows, zeroRow());
            std::uniform_real_distribution<float> u(-1.0f, 1.0f);
            if (down) {
                for (uint32_t r = p.numNodes; r < numRows; r++)
                    for (int k = 0; k < K; k++) seeded[r][k] = u(rng);
            } else {
                for (uint32_t r = 0; r < lg.N; r++)
                    for (int k = 0; k < K; k++) seeded[r][k] = u(rng);
            }
            for (int trial = 0; trial < 4; trial++) {
                GroupLimits lim;
                lim.ldsBytes = 256U << (rng() % 9);        // 256 B .. 64 KB
                lim.maxEntries = 4U << (rng() % 13);       // 4 .. 16384
                lim.maxBlocks = 1 + (uint32_t)(rng() % 600);
                lim.maxScanEntries = rng() % 3 ? (1ULL << 26) : (uint64_t)(rng() % 400);
                lim.maxLevelRows = rng() % 3 ? 0xffffffffU : 1 + (uint32_t)(rng() % 200);
                lim.classMask = rng() % 4 ? (1U << kGroupClasses) - 1 : (uint32_t)(rng() % (1U << kGroupClasses));
                const uint32_t heavy = 2 + (uint32_t)(rng() % 70);
                const std::vector<GroupPlan> plans = buildGroupPlans(p, groupRegionEnd(p, down), heavy, lim);
                if (plans.size() != kGroupClasses) throw std::runtime_error("class count");
                std::vector<Row> ref = seeded;
                runLevels(ref, p, p.firstPulledLevel, p.numLevels, heavy, start);
                for (uint32_t c = 0; c < kGroupClasses; c++) {
                    const GroupPlan& gp = plans[c];
                    if (gp.slotBytes != groupClassBytes(c) || gp.heavyFanin != heavy) throw std::runtime_error("plan header");
                    if (gp.empty()) continue;
                    validateGroupPlan(p, gp, lim);
                    if (gp.ldsBytes > lim.ldsBytes) throw std::runtime_error("LDS over the limit");
                    std::vector<Row> rows = seeded;
                    runLevels(rows, p, p.firstPulledLevel, gp.levelBegin, heavy, start);
                    runGroups(rows, gp);
                    runLevels(rows, p, gp.levelEnd, p.numLevels, heavy, start);
                    for (uint32_t r = 0; r < p.numNodes; r++)
                        if (std::memcmp(rows[r].data(), ref[r].data(), sizeof(float) * K) != 0) {
                            std::fprintf(stderr, "graph %d dir %d class %u: row %u differs\n", it, dir, c, r);
                            return 1;
                        }
                    plansChecked++;
                }
            }
        }
    }
    if (plansChecked < (uint64_t)graphs) {
        std::fprintf(stderr, "too few plans checked: %llu\n", (unsigned long long)plansChecked);
        return 1;
    }
    std::printf("grouping fuzz ok: %llu plans\n", (unsigned long long)plansChecked);
    return 0;
}
