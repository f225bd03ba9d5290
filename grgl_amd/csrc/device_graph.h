// Device-resident GRG and the traversal driver shared by the C ABI and the server.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "grgl_hip.h"
#include "levelize.h"
#include "stacking.h"

namespace grgl_amd {

struct MisuseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NoMemError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct UnsupportedError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define GRGL_HIP_CHECK(expr)                                                                           \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            if (_e == hipErrorOutOfMemory) throw ::grgl_amd::NoMemError(std::string(#expr) + ": out of device memory"); \
            throw ::grgl_amd::DeviceError(std::string(#expr) + ": " + hipGetErrorString(_e));          \
        }                                                                                              \
    } while (0)

// Arithmetic class of a dtype: signed and unsigned integers of one width share the
// wrapping add; bool is XOR on bytes.
enum TypeClass : int { kF64 = 0, kF32 = 1, kW64 = 2, kW32 = 3, kW16 = 4, kW8 = 5, kX8 = 6 };
TypeClass typeClassOf(grgl_dtype dt);
size_t dtypeSize(grgl_dtype dt);

struct RowGeometry {
    uint32_t vb;        // bytes per lane access: 4, 8 or 16
    uint32_t rowBytes;  // K * elemSize rounded up to vb
    uint64_t ld;        // row stride in bytes
    uint32_t lanes;     // rowBytes / vb
};
RowGeometry rowGeometry(uint32_t K, size_t elemSize);
// Team widths (lanes per row) of the level kernels: powers of two up to the number of
// 16-byte slices of a row rounded up, or up to 8 for rows of at most eight slices.
uint32_t maxTeamLanes(const RowGeometry& geo);
// Bytes each lane of a team of `tw` lanes accesses in the level kernels.  A team of up to 8
// lanes takes the narrowest access (4, 8 or 16 bytes, never less than one element) with
// which it covers the whole row at once, so that the row is one column slice; lanes left
// over idle, which spreads a level's rows over more wavefronts and compute units.  Rows
// too wide for that, and teams of 16 lanes or more, take 16-byte accesses.
inline uint32_t levelVb(const RowGeometry& geo, uint32_t tw, size_t elemSize) {
    if (tw > 8) return geo.vb;
    for (uint32_t v = elemSize > 4 ? (uint32_t)elemSize : 4; v < geo.vb; v *= 2)
        if (geo.rowBytes / v <= tw) return v;
    return geo.vb;
}
// Column slices (grid y) a row is cut into for a team of `tw` lanes.
inline uint32_t levelTiles(const RowGeometry& geo, uint32_t tw, size_t elemSize) {
    return (geo.rowBytes / levelVb(geo, tw, elemSize) + tw - 1) / tw;
}

// A launch's share of a level: consecutive rows whose gather lists hold at most
// kChunkEntries entries (or one longer list), so that list offsets relative to the chunk's
// first entry fit 32 bits whatever the graph's total.  Graphs under 2^30 entries per level
// have one chunk per level.
struct LevelChunk {
    uint32_t rowBegin, rowEnd;
    uint32_t heavy;       // heavy rows at the front of the chunk (heavyFanin)
    uint64_t entryBegin;  // first gather entry of rowBegin
};
constexpr uint64_t kChunkEntries = 1ULL << 30;
// Device words a chained traversal keeps its ticket and completion counters in (zero whenever
// no chained traversal runs on them).
constexpr size_t kChainStateBytes = 4096;
// Captured level sequences kept per graph before idle ones are dropped.
constexpr size_t kMaxCapturedGraphs = 64;

// Launch table of one direction for one heavy threshold and launch size: which rows at the
// front of each level are heavy, and how each level is cut into launches.  Immutable once
// built: traversals take a reference-counted snapshot, so retuning a direction swaps in a new
// table and never changes one that a traversal on another thread is reading.
// Heavy threshold of a new graph's launch tables (and of its grouped-launch plans).
constexpr uint32_t kDefaultHeavyFanin = 64;

struct LaunchTable {
    uint32_t heavyFanin = kDefaultHeavyFanin;
    uint64_t chunkEntries = kChunkEntries; // gather entries per launch at most
    std::vector<uint32_t> heavyCount;      // per level, rows with more than heavyFanin entries
    std::vector<uint64_t> levelEntries;    // per level, gather entries
    std::vector<uint64_t> heavyEntries;    // per level, gather entries of the heavy rows
    std::vector<uint64_t> levelEntryBegin; // numLevels + 1: first gather entry of each level
    std::vector<LevelChunk> chunks;        // launches of every level, in level order
    std::vector<uint32_t> levelChunk;      // numLevels + 1 offsets into chunks
};

// A grouped-launch plan (grouping.h) on the device: its arrays in one allocation.
struct DeviceGroupPlan {
    uint32_t slotBytes = 0, heavyFanin = 0, levelBegin = 0, levelEnd = 0, blocks = 0;
    uint32_t ldsBytes = 0;  // LDS of the largest block
    uint32_t numRows = 0;   // rows of the grouped levels
    uint32_t numHeavy = 0;  // heavy rows among them (64 partial prefix sums each in the row buffer's tail)
    uint64_t prefixEntries = 0; // leading list entries that read HBM rows, over all rows
    uint32_t* mem = nullptr;
    const uint32_t *blockRow = nullptr, *blockSeg = nullptr, *blockHeavy = nullptr, *blockImport = nullptr,
                   *segRow = nullptr, *segHeavy = nullptr, *segHeavyOrd = nullptr, *row = nullptr, *off = nullptr,
                   *pre = nullptr, *src = nullptr, *remOff = nullptr, *rem = nullptr, *heavyRow = nullptr,
                   *importRow = nullptr;
    bool empty() const { return levelBegin >= levelEnd; }
};
constexpr uint32_t kDeviceGroupClasses = 5; // slots of 4, 8, 16, 32, 64 bytes
constexpr uint32_t kLateLevelRows = 8192;

struct DeviceDirPlan {
    bool down = false;             // the DOWN plan (else UP)
    bool samplesLast = false;      // DOWN: every sample's row is in the last level
    bool firstLevelSeedsOnly = false; // the first pulled level's lists read only seed rows
    uint32_t numNodes = 0, numLevels = 0, firstPulledLevel = 0, maxFanIn = 0;
    uint64_t numEntries = 0;
    std::vector<uint32_t> levelStart;
    std::vector<uint32_t> fanIn;   // host: per row list length (heavy split points)
    uint32_t* off = nullptr;       // list offsets modulo 2^32 (kernels subtract a chunk's base)
    uint32_t* src = nullptr;
    uint32_t* order = nullptr;     // row -> node
    uint32_t* rank = nullptr;      // node -> row
    uint32_t* outRow = nullptr;
    uint32_t* missRow = nullptr;
    uint32_t* coalByRow = nullptr;
    // Grouped-launch plans per row class (built with the default heavy threshold; used only
    // while the launch table has that threshold).  Immutable after graph creation.
    DeviceGroupPlan groups[kDeviceGroupClasses];
    // DOWN: 64-byte slots over only the small levels before the samples (levels of at most
    // kLateLevelRows rows), for rows wider than 64 bytes run as column slices.
    DeviceGroupPlan lateWide;
    // Current launch table; replaced (never modified) under grgl_graph::mu.
    std::shared_ptr<const LaunchTable> table;
    std::shared_ptr<const LaunchTable> buildTable(uint32_t heavyFanin, uint64_t chunkEntries) const;
};

struct Workspace {
    char* raw = nullptr;        // the allocation: chain state area, then `bytes` at ptr
    char* ptr = nullptr;        // row buffer (and host-buffer staging)
    uint32_t* chainState = nullptr; // kChainStateBytes, zero whenever no traversal runs
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    bool used = false;
};

struct Tuning {
    uint32_t teamLanes = 0; // 0 = default
    uint32_t useGraph = 0;
    uint32_t xcdMap = 0;    // 0 = default (off), 1 = XCD-aware block order, 2 = off
    uint32_t chain = 0;     // 0 = default, 1 = level-by-level launches, 2 = chained when eligible,
                            // 3 = grouped levels when eligible
    uint32_t blockThreads = 0;  // wide-team level kernels: 0 = default (256), 256 or 512 threads
    uint32_t loadsInFlight = 0; // wide-team level kernels: 0 = default (4), 4 or 8 row loads per batch
};

// Host-mapped control words of a graph's chained traversals (page-locked, device-visible).
struct ChainControl {
    uint32_t abortWord; // nonzero: chained traversals give up at their next wait
    uint32_t gaveUp;    // 1: a chained traversal gave up (its outputs are undefined)
};

// A level sequence as an instantiated hipGraph.  Held by shared_ptr: a traversal keeps its reference from the
// cache lookup until its launch is recorded, and the cache drops only entries that nothing
// else holds and whose last launch has completed.
struct CapturedLevels {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t lastLaunch = nullptr; // recorded after every launch of `exec`
    CapturedLevels() = default;
    CapturedLevels(const CapturedLevels&) = delete;
    CapturedLevels& operator=(const CapturedLevels&) = delete;
    ~CapturedLevels() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (lastLaunch) (void)hipEventDestroy(lastLaunch);
    }
};

} // namespace grgl_amd

struct grgl_graph {
    int device = 0;
    uint32_t N = 0, V = 0, M = 0, P = 1;
    uint64_t E = 0;
    bool hasMiss = false, hasCoal = false;
    grgl_amd::DeviceDirPlan up, down;
    uint64_t deviceBytes = 0;
    grgl_amd::ChainControl* chainCtl = nullptr;  // host pointer (hipHostMalloc, mapped)
    grgl_amd::ChainControl* chainCtlDev = nullptr; // its device address
    std::atomic<uint32_t> chainWaitUs{200000};  // longest wait of a chained traversal, us
    uint64_t wallClockKHz = 100000;              // rate of the device clock the waits read
    std::mutex mu;
    std::vector<std::unique_ptr<grgl_amd::Workspace>> workspaces; // stable addresses
    std::vector<hipStream_t> freeStreams;
    std::map<std::tuple<int, int, uint32_t>, grgl_amd::Tuning> tuning; // (dir, typeClass, K)
    // Kernel timing (level kernels); read by traversals on any thread.
    std::atomic<bool> timing{false};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timingEvents;
    std::vector<double> timingBytes;
    double timedMs = 0.0;
    uint64_t timedLaunches = 0;
    double timedBytes = 0.0;
    // Captured level sequences: (dir, typeClass, vb, tw, ld, rowBytes, flags and heavy threshold,
    // rows ptr, launch size, level-sequence form)
    std::map<std::tuple<int, int, uint32_t, uint32_t, uint64_t, uint32_t, int, uintptr_t, uint64_t, uint32_t>,
             std::shared_ptr<grgl_amd::CapturedLevels>>
        graphs;
};

namespace grgl_amd {

struct TravRequest {
    uint32_t K;
    const void* in;   // device, K x inCols
    void* out;        // device, K x outCols
    uint32_t kOffset;
};

struct TravOptions {
    bool emitAll = false;
    bool byIndividual = false;
    grgl_init_mode initMode = GRGL_INIT_ZERO;
    const void* init = nullptr;  // device
    void* miss = nullptr;        // device
    bool accumulate = false;
    bool orderPreserving = false;
    // Direct kernel launches only: no hipGraph, no timing events (the allocation-free entry
    // point, whose callers may not allocate).
    bool launchOnly = false;
    // kChainStateBytes of zeroed device memory owned by the caller's buffer, or null: chained
    // (single-launch) level sequences are possible only with it.
    uint32_t* chainState = nullptr;
};

// Size in bytes of the row buffer for a traversal of Ktotal stacked rows.
size_t rowsBufferBytes(const grgl_graph* g, grgl_direction dir, uint32_t Ktotal, size_t elemSize);

// Enqueue one traversal on `stream`: seed, all levels, read-out for every request.
// `rows` must hold rowsBufferBytes(...) bytes.  Device pointers only.
void enqueueTraversal(grgl_graph* g, grgl_dtype dt, grgl_direction dir, const std::vector<TravRequest>& reqs,
                      uint32_t Ktotal, const TravOptions& opt, char* rows, hipStream_t stream);

// Workspace cache (stream-ordered reuse).
Workspace* acquireWorkspace(grgl_graph* g, size_t bytes, hipStream_t stream);
void releaseWorkspace(grgl_graph* g, Workspace* ws, hipStream_t stream);
hipStream_t acquireInternalStream(grgl_graph* g);
void releaseInternalStream(grgl_graph* g, hipStream_t s);

void setLastError(const std::string& msg);

// All argument checks of grgl_hip_matmul (internal_mult.h:130-158, _grgl.cpp:100-185).
void validateMatmulArgs(const grgl_graph* g, grgl_dtype dtype, const void* input, size_t input_cols, size_t input_rows,
                        grgl_direction direction, const void* output, size_t output_size, int emit_all_nodes,
                        int by_individual, const void* init, grgl_init_mode init_mode, const void* miss);

// Status code of an exception caught at the C ABI; sets the thread-local message.
grgl_status statusOf(const std::exception& e);

// Validation shared by grgl_hip_matmul and the server (internal_mult.h:130-158).
void validateShapes(const grgl_graph* g, grgl_direction dir, size_t inCols, size_t rows, size_t outSize,
                    bool emitAll, bool byIndividual);

} // namespace grgl_amd
