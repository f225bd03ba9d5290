// Levelized-graph files under damage (run by tests/test_levelized_fuzz.py, built with the
// address and undefined-behaviour sanitizers): random ordered graphs are levelized, written,
// read back equal, and then
//   * plans with one entry or size changed must either be rejected by validateLevelized or
//     still pass it (a harmless change) — never read out of bounds while being checked;
//   * files with flipped bytes or cut short must be rejected by readLevelizedFile.
// Exit code 0 = all checks held.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "levelize.h"

using namespace grgl_amd;

static HostGrg randomGraph(std::mt19937_64& rng) {
    HostGrg g;
    const uint32_t ploidy = 2;
    const uint32_t N = 2 * (2 + (uint32_t)(rng() % 20));
    const uint32_t internal = 1 + (uint32_t)(rng() % 60);
    g.numSamples = N;
    g.ploidy = ploidy;
    g.downOff.assign(1, 0);
    for (uint32_t s = 0; s < N; s++) g.downOff.push_back(0);
    for (uint32_t v = N; v < N + internal; v++) {
        std::vector<uint32_t> kids;
        const uint32_t want = 1 + (uint32_t)(rng() % (rng() % 4 == 0 ? 90 : 5));
        for (uint32_t i = 0; i < want; i++) kids.push_back((uint32_t)(rng() % v));
        std::sort(kids.begin(), kids.end());
        kids.erase(std::unique(kids.begin(), kids.end()), kids.end());
        g.downChild.insert(g.downChild.end(), kids.begin(), kids.end());
        g.downOff.push_back(g.downChild.size());
    }
    const uint32_t V = N + internal;
    const uint32_t M = (uint32_t)(rng() % 80);
    for (uint32_t m = 0; m < M; m++) g.mutNode.push_back(rng() % 7 == 0 ? kInvalidNode : (uint32_t)(rng() % V));
    if (rng() % 2) {
        for (uint32_t m = 0; m < M; m++) g.mutMiss.push_back(rng() % 3 ? kInvalidNode : (uint32_t)(rng() % V));
    }
    if (rng() % 2) {
        g.coal.assign(V, 0);
        for (uint32_t v = N; v < V; v++) g.coal[v] = (uint32_t)(rng() % 5);
    }
    for (uint32_t m = 0; m < M; m++) g.mutPos.push_back((double)m);
    return g;
}

static std::vector<char> slurp(const std::string& p) {
    std::ifstream f(p, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

static void spill(const std::string& p, const std::vector<char>& b, size_t n) {
    std::ofstream f(p, std::ios::binary | std::ios::trunc);
    f.write(b.data(), (std::streamsize)n);
}

static std::vector<uint32_t>* pickArray(DirPlan& p, unsigned k) {
    switch (k % 8) {
    case 0: return &p.levelStart;
    case 1: return &p.order;
    case 2: return &p.rank;
    case 3: return &p.off;
    case 4: return &p.src;
    case 5: return &p.outRow;
    case 6: return &p.missRow;
    default: return &p.coalByRow;
    }
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int graphs = argc > 2 ? std::atoi(argv[2]) : 60;
    const int perGraph = argc > 3 ? std::atoi(argv[3]) : 200;
    std::mt19937_64 rng(20261204);
    long rejected = 0, accepted = 0, fileRejected = 0;
    int failures = 0;
    const std::string path = dir + "/fuzz.lvlz", bad = dir + "/fuzz_bad.lvlz";
    for (int gi = 0; gi < graphs; gi++) {
        HostGrg g = randomGraph(rng);
        LevelizedGraph lg = levelize(g);
        validateLevelized(lg);
        writeLevelizedFile(lg, path);
        LevelizedGraph back = readLevelizedFile(path);
        if (back.up.src != lg.up.src || back.down.order != lg.down.order || back.up.coalByRow != lg.up.coalByRow ||
            back.down.outRow != lg.down.outRow || back.up.missRow != lg.up.missRow) {
            std::printf("FAILED: round trip differs (graph %d)\n", gi);
            failures++;
        }
        const uint32_t V = lg.V, M = lg.M;
        const uint32_t special[] = {0, 1, V - 1, V, V + 1, V + M, V + 2 * M, 0x7fffffffU, 0xffffffffU, 0x80000000U};
        for (int it = 0; it < perGraph; it++) {
            LevelizedGraph m = lg;
            DirPlan& p = (rng() % 2) ? m.up : m.down;
            const unsigned what = (unsigned)(rng() % 11);
            if (what < 8) {
                std::vector<uint32_t>* a = pickArray(p, what);
                if (a->empty() || rng() % 10 == 0) {
                    a->push_back((uint32_t)rng()); // size change
                } else {
                    (*a)[rng() % a->size()] = (rng() % 2) ? special[rng() % 10] : (uint32_t)(rng() % (V + 2 * M + 2));
                }
            } else if (what == 8) {
                p.numLevels += (rng() % 2) ? 1 : (uint32_t)-1;
            } else if (what == 9) {
                p.maxFanIn = (uint32_t)(rng() % 100);
            } else {
                p.firstPulledLevel ^= 1;
            }
            try {
                validateLevelized(m);
                accepted++;
            } catch (const std::runtime_error&) {
                rejected++;
            }
        }
        const std::vector<char> bytes = slurp(path);
        for (int it = 0; it < 20; it++) {
            std::vector<char> b = bytes;
            const int flips = 1 + (int)(rng() % 3);
            for (int f = 0; f < flips; f++) b[rng() % b.size()] ^= (char)(1 + rng() % 255);
            const size_t n = (it % 4 == 0) ? (size_t)(rng() % b.size()) : b.size();
            spill(bad, b, n);
            try {
                readLevelizedFile(bad);
                std::printf("FAILED: damaged file accepted (graph %d)\n", gi);
                failures++;
            } catch (const std::runtime_error&) {
                fileRejected++;
            }
        }
    }
    std::printf("plans rejected %ld, accepted %ld; damaged files rejected %ld\n", rejected, accepted, fileRejected);
    if (rejected == 0) failures++;
    if (!failures) std::printf("levelized fuzz ok\n");
    return failures ? 1 : 0;
}
