// CPU emulation of the chained traversal's claim / wait protocol (pull_chain_kernel in
// grgl_amd/csrc/kernels.h), with threads standing for thread blocks:
//   * tasks are numbered level by level; a thread claims the next task number from a ticket
//     counter and finishes it before claiming another;
//   * a task of level l > 0 waits (bounded in time, polling the give-up flag and the abort word)
//     until all tasks of level l - 1 are counted complete, then reads rows of lower levels,
//     writes its own rows and counts itself complete with a release increment;
//   * a wait that gives up sets the flag; threads stop claiming when the flag is set;
//   * the last thread to leave counts a give-up and clears the state words.
// Checks: results equal a sequential level-by-level evaluation for every thread count (1 to
// 96 threads, fewer or more than the tasks), with random delays; the state is zero afterwards;
// with the abort word set every thread leaves, the give-up is counted once and the state is
// cleared; a too-short wait bound gives up instead of hanging.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

namespace {

struct Graph {
    std::vector<uint32_t> levelStart;           // rows of each level
    std::vector<std::vector<uint32_t>> sources; // per row: rows of lower levels it sums
};

// Random levelled DAG: level 0 holds `base` rows; a row of level l sums 1..5 rows of levels < l,
// at least one from level l - 1.
Graph makeGraph(uint32_t levels, uint32_t base, uint32_t seed) {
    std::mt19937 rng(seed);
    Graph g;
    g.levelStart.push_back(0);
    uint32_t rows = 0;
    for (uint32_t l = 0; l < levels; l++) {
        const uint32_t n = l == 0 ? base : 1 + rng() % (2 * base);
        for (uint32_t i = 0; i < n; i++) {
            std::vector<uint32_t> src;
            if (l > 0) {
                const uint32_t prevLo = g.levelStart[l - 1], prevHi = g.levelStart[l];
                src.push_back(prevLo + rng() % (prevHi - prevLo));
                const uint32_t extra = rng() % 5;
                for (uint32_t k = 0; k < extra; k++) src.push_back(rng() % prevHi);
            }
            g.sources.push_back(src);
        }
        rows += n;
        g.levelStart.push_back(rows);
    }
    return g;
}

struct Chain {
    const Graph* g;
    uint32_t rowsPerTask;
    std::vector<uint32_t> taskStart; // per pulled level (levels 1..), plus the total
    std::atomic<uint32_t> ticket{0}, exits{0}, flag{0};
    std::vector<std::atomic<uint32_t>> done;
    std::atomic<uint32_t>* abortWord;
    std::atomic<uint32_t>* gaveUp;
    std::chrono::microseconds waitBound;
    std::vector<uint64_t> value; // row values; level-0 rows are seeds
    Chain(const Graph* graph, uint32_t rpt, std::atomic<uint32_t>* abortW, std::atomic<uint32_t>* gu,
          std::chrono::microseconds bound)
        : g(graph), rowsPerTask(rpt), done(graph->levelStart.size()), abortWord(abortW), gaveUp(gu), waitBound(bound) {
        const uint32_t L = (uint32_t)graph->levelStart.size() - 1;
        uint32_t t = 0;
        for (uint32_t l = 1; l < L; l++) {
            taskStart.push_back(t);
            t += (graph->levelStart[l + 1] - graph->levelStart[l] + rpt - 1) / rpt;
        }
        taskStart.push_back(t);
        value.assign(graph->levelStart.back(), 0);
        for (uint32_t r = 0; r < graph->levelStart[1]; r++) value[r] = 1000 + r;
    }
    uint32_t numLevels() const { return (uint32_t)taskStart.size() - 1; }
    uint32_t total() const { return taskStart.back(); }
};

void block(Chain& c, uint32_t blocks, uint32_t seed) {
    std::mt19937 rng(seed);
    auto jitter = [&] {
        if (rng() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 50));
    };
    if (c.abortWord->load(std::memory_order_relaxed) != 0) c.flag.store(1, std::memory_order_relaxed);
    for (;;) {
        uint32_t task = c.total();
        if (c.flag.load(std::memory_order_relaxed) == 0) task = c.ticket.fetch_add(1, std::memory_order_relaxed);
        if (task >= c.total()) break;
        uint32_t lo = 0, hi = c.numLevels() - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) / 2;
            if (c.taskStart[mid] <= task) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t l = lo; // pulled level index; graph level l + 1
        jitter();              // structure reads before the wait
        bool ok = true;
        if (l > 0) {
            const uint32_t need = c.taskStart[l] - c.taskStart[l - 1];
            const auto t0 = std::chrono::steady_clock::now();
            for (uint32_t spin = 0; c.done[l - 1].load(std::memory_order_relaxed) < need; spin++) {
                if ((spin & 63) == 63 && (c.flag.load(std::memory_order_relaxed) != 0 ||
                                          c.abortWord->load(std::memory_order_relaxed) != 0 ||
                                          std::chrono::steady_clock::now() - t0 > c.waitBound)) {
                    ok = false;
                    break;
                }
                std::this_thread::yield();
            }
            // The kernel polls with relaxed loads and then performs an acquire fence; here the
            // acquire is an acquire load of the counter (equivalent for this pattern, and
            // ThreadSanitizer, which does not model standalone fences, can check it).
            if (ok) (void)c.done[l - 1].load(std::memory_order_acquire);
        }
        if (!ok) {
            c.flag.store(1, std::memory_order_relaxed);
            break;
        }
        const uint32_t lev = l + 1;
        const uint32_t r0 = c.g->levelStart[lev] + (task - c.taskStart[l]) * c.rowsPerTask;
        const uint32_t r1 = std::min<uint32_t>(r0 + c.rowsPerTask, c.g->levelStart[lev + 1]);
        for (uint32_t r = r0; r < r1; r++) {
            uint64_t s = 0;
            for (uint32_t src : c.g->sources[r]) s += c.value[src] * 3 + 1;
            c.value[r] = s;
        }
        jitter();
        c.done[l].fetch_add(1, std::memory_order_release);
    }
    const uint32_t out = c.exits.fetch_add(1, std::memory_order_acq_rel);
    if (out == blocks - 1) {
        if (c.flag.load(std::memory_order_relaxed) != 0) c.gaveUp->fetch_add(1, std::memory_order_relaxed);
        c.ticket.store(0, std::memory_order_relaxed);
        c.exits.store(0, std::memory_order_relaxed);
        c.flag.store(0, std::memory_order_relaxed);
        for (auto& d : c.done) d.store(0, std::memory_order_relaxed);
    }
}

std::vector<uint64_t> sequential(const Graph& g) {
    std::vector<uint64_t> v(g.levelStart.back(), 0);
    for (uint32_t r = 0; r < g.levelStart[1]; r++) v[r] = 1000 + r;
    for (uint32_t r = g.levelStart[1]; r < g.levelStart.back(); r++) {
        uint64_t s = 0;
        for (uint32_t src : g.sources[r]) s += v[src] * 3 + 1;
        v[r] = s;
    }
    return v;
}

bool stateClear(Chain& c) {
    if (c.ticket.load() || c.exits.load() || c.flag.load()) return false;
    for (auto& d : c.done)
        if (d.load()) return false;
    return true;
}

uint32_t run(Chain& c, uint32_t blocks) {
    std::vector<std::thread> th;
    for (uint32_t b = 0; b < blocks; b++) th.emplace_back(block, std::ref(c), blocks, 7919 * b + 1);
    for (auto& t : th) t.join();
    return c.gaveUp->load();
}

} // namespace

int main() {
    int failures = 0;
    for (uint32_t seed = 1; seed <= 6; seed++) {
        const Graph g = makeGraph(12 + seed * 3, 40 + 30 * seed, seed);
        const std::vector<uint64_t> want = sequential(g);
        for (uint32_t blocks : {1U, 2U, 7U, 33U, 96U}) {
            for (uint32_t rpt : {1U, 16U, 64U}) {
                std::atomic<uint32_t> abortWord{0}, gaveUp{0};
                Chain c(&g, rpt, &abortWord, &gaveUp, std::chrono::microseconds(20000000));
                for (int rep = 0; rep < 2; rep++) { // the state clears itself for the next run
                    std::fill(c.value.begin() + g.levelStart[1], c.value.end(), 0);
                    const uint32_t gu = run(c, blocks);
                    if (gu != 0 || c.value != want || !stateClear(c)) {
                        std::printf("FAIL seed %u blocks %u rows/task %u rep %d (gave up %u)\n", seed, blocks, rpt, rep,
                                    gu);
                        failures++;
                    }
                }
            }
        }
        // Abort word set: every thread leaves; one give-up is counted; the state is cleared.
        for (uint32_t blocks : {1U, 5U, 64U}) {
            std::atomic<uint32_t> abortWord{1}, gaveUp{0};
            Chain c(&g, 8, &abortWord, &gaveUp, std::chrono::microseconds(20000000));
            const uint32_t gu = run(c, blocks);
            if (gu != 1 || !stateClear(c)) {
                std::printf("FAIL abort seed %u blocks %u (gave up %u)\n", seed, blocks, gu);
                failures++;
            }
        }
    }
    // A wait bound too short for the work gives up rather than waiting: one thread holds a task
    // of level 1 while the others wait on it; the run ends either complete or counted.
    {
        const Graph g = makeGraph(30, 2000, 99);
        const std::vector<uint64_t> want = sequential(g);
        std::atomic<uint32_t> abortWord{0}, gaveUp{0};
        Chain c(&g, 1, &abortWord, &gaveUp, std::chrono::microseconds(1));
        const uint32_t gu = run(c, 32);
        if (!stateClear(c) || (gu == 0 && c.value != want) || gu > 1) {
            std::printf("FAIL short bound (gave up %u)\n", gu);
            failures++;
        }
    }
    std::printf(failures ? "chain protocol: %d failures\n" : "chain protocol ok\n", failures);
    return failures ? 1 : 0;
}
