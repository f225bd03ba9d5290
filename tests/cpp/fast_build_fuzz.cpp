// The haplotype-clustering constructor under AddressSanitizer and UndefinedBehaviorSanitizer
// (run by tests/test_fast_build_sanitized.py): random haplotype tables (copied founder blocks
// with noise, empty, full and duplicate rows, missing-data rows) built with random window
// lengths, rebuild proportions, tree splits and direct-mapping thresholds; every mutation's
// samples, read back from the finished graph, equal its carriers, each exactly once.
// Exit code 0 = all checks held.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "fast_build.h"

using namespace grgl_amd;

int main(int argc, char** argv) {
    const int cases = argc > 1 ? std::atoi(argv[1]) : 150;
    std::mt19937_64 rng(7);
    int failures = 0;
    for (int c = 0; c < cases; c++) {
        const uint32_t N = 2 * (2 + (uint32_t)(rng() % 40));
        const uint32_t M = 1 + (uint32_t)(rng() % 500);
        const uint32_t F = 1 + (uint32_t)(rng() % 6);
        std::vector<std::vector<uint8_t>> founder(F, std::vector<uint8_t>(M));
        const double density = 0.05 + (rng() % 50) / 100.0;
        for (auto& f : founder)
            for (auto& b : f) b = (rng() % 1000) < density * 1000;
        std::vector<std::vector<uint8_t>> hap(N);
        for (auto& h : hap) {
            h = founder[rng() % F];
            for (auto& b : h)
                if (rng() % 100 < 3) b ^= 1;
        }
        std::vector<MutationRecordIn> rows;
        std::vector<std::vector<uint32_t>> sets;
        std::vector<std::vector<uint32_t>> missing; // per row: missing haplotypes of its site (empty if none)
        double pos = 0;
        for (uint32_t m = 0; m < M; m++) {
            pos += 1 + rng() % 3;
            std::vector<uint32_t> miss;
            if (rng() % 25 == 0) {
                for (uint32_t s = 0; s < N; s++)
                    if (rng() % 10 == 0) miss.push_back(s);
                if (!miss.empty()) {
                    MutationRecordIn r;
                    r.position = pos;
                    r.allele = ".";
                    rows.push_back(r);
                    sets.push_back(miss);
                    missing.push_back({});
                }
            }
            const int kind = (int)(rng() % 40);
            std::vector<uint32_t> S;
            for (uint32_t s = 0; s < N; s++) {
                const bool carry = kind == 0 ? false : kind == 1 ? true : hap[s][m] != 0;
                if (carry && !std::binary_search(miss.begin(), miss.end(), s)) S.push_back(s);
            }
            MutationRecordIn r;
            r.position = pos;
            r.allele = "G";
            r.refAllele = "A";
            rows.push_back(r);
            sets.push_back(S);
            missing.push_back(miss);
            if (rng() % 30 == 0) { // a duplicate row (another allele at the site)
                r.allele = "T";
                rows.push_back(r);
                sets.push_back(S);
                missing.push_back(miss);
            }
        }
        FastBuildParams p;
        const uint32_t lengths[] = {1, 2, 5, 16, 63, 254};
        p.hapLength = lengths[rng() % 6];
        const double rebuild[] = {0.01, 0.1, 0.5, 2.0};
        p.rebuildProportion = rebuild[rng() % 4];
        p.treeCount = (rng() % 3 == 0) ? (uint32_t)(1 + rng() % 9) : 0;
        p.hapSumMultiplier = (rng() % 2) ? 1.0 : 0.5;
        p.noTreeBelow = (rng() % 3 == 0) ? (double)(rng() % 6) : 0.0;
        FastBuildStats st;
        MutableGraph g = fastBuildFromSamples(N, 2, rows, sets, p, st, 1 + (uint32_t)(rng() % 4));
        const HostGrg h = g.freeze();
        size_t k = 0;
        for (size_t i = 0; i < rows.size(); i++) {
            if (rows[i].allele == ".") continue;
            const uint32_t node = h.mutNode[k];
            std::vector<uint32_t> got;
            if (node != kInvalidNode) {
                std::vector<uint32_t> stack(1, node);
                while (!stack.empty()) {
                    const uint32_t v = stack.back();
                    stack.pop_back();
                    if (v < N) {
                        got.push_back(v);
                        continue;
                    }
                    for (uint64_t e = h.downOff[v]; e < h.downOff[v + 1]; e++) stack.push_back(h.downChild[e]);
                }
            }
            std::sort(got.begin(), got.end());
            if (got != sets[i]) {
                std::printf("FAILED: case %d row %zu: %zu samples, want %zu\n", c, i, got.size(), sets[i].size());
                failures++;
            }
            const uint32_t miss = h.mutMiss.empty() ? kInvalidNode : h.mutMiss[k];
            if ((miss == kInvalidNode) != missing[i].empty()) {
                std::printf("FAILED: case %d row %zu: missingness node %s\n", c, i,
                            miss == kInvalidNode ? "absent" : "unexpected");
                failures++;
            }
            k++;
        }
        if (k != h.numMutations()) {
            std::printf("FAILED: case %d: %zu mutations, graph has %u\n", c, k, h.numMutations());
            failures++;
        }
    }
    if (!failures) std::printf("fast build fuzz ok (%d cases)\n", cases);
    return failures ? 1 : 0;
}
