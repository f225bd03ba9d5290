// libgrgl_hip.so: device graph upload, the traversal driver and the C ABI of grgl_hip.h.
#include <functional>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <memory>

#include "device_graph.h"
#include "grouping.h"
#include "kernels.h"

#define GRGL_EXPORT extern "C" __attribute__((visibility("default")))

namespace grgl_amd {

namespace {
thread_local std::string g_lastError;

template <typename T> T* uploadVec(const std::vector<T>& v, uint64_t& bytes) {
    if (v.empty()) return nullptr;
    T* d = nullptr;
    GRGL_HIP_CHECK(hipMalloc(&d, v.size() * sizeof(T)));
    GRGL_HIP_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    bytes += v.size() * sizeof(T);
    return d;
}

void freeDev(void* p) {
    if (p) (void)hipFree(p);
}

uint32_t pow2ceil(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

uint32_t log2u(uint32_t x) {
    uint32_t l = 0;
    while ((1U << l) < x) l++;
    return l;
}


} // namespace

void setLastError(const std::string& msg) { g_lastError = msg; }

TypeClass typeClassOf(grgl_dtype dt) {
    switch (dt) {
    case GRGL_F64: return kF64;
    case GRGL_F32: return kF32;
    case GRGL_I64:
    case GRGL_U64: return kW64;
    case GRGL_I32:
    case GRGL_U32: return kW32;
    case GRGL_I16:
    case GRGL_U16: return kW16;
    case GRGL_I8:
    case GRGL_U8: return kW8;
    case GRGL_BOOL: return kX8;
    }
    throw MisuseError("Unsupported dtype code " + std::to_string((int)dt));
}

size_t dtypeSize(grgl_dtype dt) {
    switch (typeClassOf(dt)) {
    case kF64:
    case kW64: return 8;
    case kF32:
    case kW32: return 4;
    case kW16: return 2;
    default: return 1;
    }
}

RowGeometry rowGeometry(uint32_t K, size_t elemSize) {
    RowGeometry r;
    const uint64_t raw = (uint64_t)K * elemSize;
    r.vb = raw <= 4 ? 4 : (raw <= 8 ? 8 : 16);
    const uint64_t rb = (raw + r.vb - 1) / r.vb * r.vb;
    if (rb > 0xffffffffULL) throw MisuseError("too many rows (K) for one traversal");
    r.rowBytes = (uint32_t)rb;
    // Row starts aligned so that the 256-byte slices a team reads never straddle an extra
    // cache line: 256-byte alignment for wide rows, 64 bytes for medium ones; rows under
    // 64 bytes (4, 8, 16, 32 or 48 bytes) get a power-of-two stride, so none straddles a
    // 64-byte boundary (48-byte rows at a 48-byte stride did, and ran 1.5x slower).
    r.ld = rb >= 256 ? (rb + 255) / 256 * 256 : (rb >= 64 ? (rb + 63) / 64 * 64 : pow2ceil((uint32_t)rb));
    r.lanes = r.rowBytes / r.vb;
    return r;
}

uint32_t maxTeamLanes(const RowGeometry& geo) {
    return std::min<uint32_t>(64, std::max<uint32_t>(8, pow2ceil(geo.lanes)));
}

std::shared_ptr<const LaunchTable> DeviceDirPlan::buildTable(uint32_t heavyFanin, uint64_t chunkEntries) const {
    auto t = std::make_shared<LaunchTable>();
    t->heavyFanin = heavyFanin;
    t->chunkEntries = chunkEntries;
    t->heavyCount.assign(numLevels, 0);
    t->levelEntries.assign(numLevels, 0);
    t->heavyEntries.assign(numLevels, 0);
    t->levelChunk.assign(numLevels + 1, 0);
    t->levelEntryBegin.assign(numLevels + 1, 0);
    std::vector<LevelChunk>& chunks = t->chunks;
    uint64_t entry = 0; // first entry of level l
    for (uint32_t l = 0; l < numLevels; l++) {
        t->levelChunk[l] = (uint32_t)chunks.size();
        t->levelEntryBegin[l] = entry;
        uint32_t lo = levelStart[l], hi = levelStart[l + 1];
        for (uint32_t r = lo; r < hi; r++) t->levelEntries[l] += fanIn[r];
        const uint32_t begin = lo;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (fanIn[mid] > heavyFanin) lo = mid + 1;
            else hi = mid;
        }
        const uint32_t heavyEnd = lo;
        t->heavyCount[l] = heavyEnd - begin;
        for (uint32_t r = begin; r < heavyEnd; r++) t->heavyEntries[l] += fanIn[r];
        if (l >= firstPulledLevel && levelStart[l + 1] > begin) {
            // Cut the level into chunks of at most chunkEntries entries (a longer list
            // gets a chunk of its own); one chunk unless the level is that large.
            LevelChunk c{begin, begin, 0, entry};
            uint64_t inChunk = 0, e = entry;
            for (uint32_t r = begin; r < levelStart[l + 1]; r++) {
                if (r > c.rowBegin && inChunk + fanIn[r] > chunkEntries) {
                    c.rowEnd = r;
                    chunks.push_back(c);
                    c = LevelChunk{r, r, 0, e};
                    inChunk = 0;
                }
                inChunk += fanIn[r];
                e += fanIn[r];
            }
            c.rowEnd = levelStart[l + 1];
            chunks.push_back(c);
            for (uint32_t k = t->levelChunk[l]; k < chunks.size(); k++) {
                LevelChunk& ck = chunks[k];
                ck.heavy = heavyEnd > ck.rowBegin ? std::min(heavyEnd, ck.rowEnd) - ck.rowBegin : 0;
            }
        }
        entry += t->levelEntries[l];
    }
    t->levelChunk[numLevels] = (uint32_t)chunks.size();
    t->levelEntryBegin[numLevels] = entry;
    return t;
}

// Snapshot of a direction's current launch table.
std::shared_ptr<const LaunchTable> launchTable(grgl_graph* g, grgl_direction dir) {
    std::lock_guard<std::mutex> lk(g->mu);
    return (dir == GRGL_DIRECTION_UP ? g->up : g->down).table;
}

// Byte offset of the grouped launches' scratch area in a row buffer: after the node and seed
// rows, 256-byte aligned.
size_t groupScratchOffset(const grgl_graph* g, grgl_direction dir, const RowGeometry& geo) {
    uint64_t rows = g->V;
    if (dir == GRGL_DIRECTION_DOWN) rows += (uint64_t)g->M * (g->hasMiss ? 2 : 1);
    return (size_t)((rows * geo.ld + 255) / 256 * 256);
}

size_t rowsBufferBytes(const grgl_graph* g, grgl_direction dir, uint32_t Ktotal, size_t elemSize) {
    const RowGeometry geo = rowGeometry(Ktotal, elemSize);
    size_t bytes = groupScratchOffset(g, dir, geo);
    // Rows of at most 8 bytes may take grouped levels with a prefix launch: 64 partial sums per
    // heavy grouped row.
    if (geo.rowBytes <= 8) {
        uint32_t c = 0;
        while ((4U << c) < geo.rowBytes) c++;
        const DeviceGroupPlan& gp = (dir == GRGL_DIRECTION_UP ? g->up : g->down).groups[c];
        if (!gp.empty()) bytes += (size_t)gp.numHeavy * 64 * gp.slotBytes;
    }
    return bytes + 256;
}

void validateShapes(const grgl_graph* g, grgl_direction dir, size_t inCols, size_t rows, size_t outSize,
                    bool emitAll, bool byIndividual) {
    // release_asserts of grg.h:1353-1354 and internal_mult.h:139 become misuse errors here.
    if (rows == 0 || inCols == 0) throw MisuseError("matmul() requires non-zero dimensions.");
    if (outSize % rows != 0) throw MisuseError("Output size is not a multiple of the number of input rows");
    const size_t outCols = outSize / rows;
    const size_t expectSampleCols = byIndividual ? g->N / g->P : g->N;
    if (dir == GRGL_DIRECTION_DOWN) {
        if (inCols != g->M) throw MisuseError("Input vector is not size M (numMutations)");
        if (!emitAll && outCols != expectSampleCols)
            throw MisuseError("Output vector is not size N (numSamples, or numIndividuals depending on flags)");
    } else if (dir == GRGL_DIRECTION_UP) {
        if (inCols != expectSampleCols)
            throw MisuseError("Input vector is not size N (numSamples, or numIndividuals depending on flags)");
        if (!emitAll && outCols != g->M) throw MisuseError("Output vector is not size M (numMutations)");
    } else {
        throw MisuseError("Invalid traversal direction");
    }
    if (emitAll && outCols != g->V) throw MisuseError("Output vector is not size numNodes");
}

// ---------------------------------------------------------------------------------------
// Kernel dispatch
// ---------------------------------------------------------------------------------------
namespace {

// Accesses narrower than 16 bytes come with teams of at most 8 lanes (levelVb).  h32: the
// variant with per-wavefront heavy rows of 32 bytes, for (vb, tw) of a 32-byte row.  The
// level kernels are resolved to their addresses, so that one launch list serves direct
// launches (hipLaunchKernel) and graph nodes (hipGraphAddKernelNode) alike.
// Wide-team shapes (16-byte accesses, teams of 8-32 lanes) of the 4- and 8-byte types also come
// with 512-thread blocks and 8 row loads in flight per batch (tuning dimensions).
template <typename T, bool XOR> constexpr bool kTileTunable = !XOR && sizeof(T) >= 4;
inline bool tileTunable(uint32_t vb, uint32_t tw) { return vb == 16 && tw >= 8 && tw <= 32; }

template <typename T, bool XOR, int TW> const void* wideKernel(uint32_t blk, uint32_t unr) {
    if constexpr (kTileTunable<T, XOR>) {
        if (blk == 512 && unr == 8) return reinterpret_cast<const void*>(&pull_level_kernel<T, 16, TW, XOR, false, 512, 8>);
        if (blk == 512) return reinterpret_cast<const void*>(&pull_level_kernel<T, 16, TW, XOR, false, 512, 4>);
        if (unr == 8) return reinterpret_cast<const void*>(&pull_level_kernel<T, 16, TW, XOR, false, kBlock, 8>);
    }
    return reinterpret_cast<const void*>(&pull_level_kernel<T, 16, TW, XOR>);
}

template <typename T, bool XOR>
const void* pullKernelT(uint32_t vb, uint32_t tw, bool h32, uint32_t blk, uint32_t unr) {
#define GRGL_PULL(VB, TW) return reinterpret_cast<const void*>(&pull_level_kernel<T, VB, TW, XOR>)
#define GRGL_PULL32(VB, TW) return reinterpret_cast<const void*>(&pull_level_kernel<T, VB, TW, XOR, true>)
    if (h32) {
        if (vb == 16 && tw == 1) GRGL_PULL32(16, 1);
        if (vb == 16 && tw == 2) GRGL_PULL32(16, 2);
        if (vb == 8 && tw == 4) GRGL_PULL32(8, 4);
        if (vb == 8 && tw == 8) GRGL_PULL32(8, 8);
        if constexpr (sizeof(T) <= 4) {
            if (vb == 4 && tw == 8) GRGL_PULL32(4, 8);
        }
        throw std::logic_error("no 32-byte heavy-row variant for this team shape");
    }
#undef GRGL_PULL32
    if (vb == 4) {
        if constexpr (sizeof(T) <= 4) {
            if (tw == 1) GRGL_PULL(4, 1);
            if (tw == 2) GRGL_PULL(4, 2);
            if (tw == 4) GRGL_PULL(4, 4);
            GRGL_PULL(4, 8);
        }
    } else if (vb == 8) {
        if constexpr (sizeof(T) <= 8) {
            if (tw == 1) GRGL_PULL(8, 1);
            if (tw == 2) GRGL_PULL(8, 2);
            if (tw == 4) GRGL_PULL(8, 4);
            GRGL_PULL(8, 8);
        }
    } else {
        switch (tw) {
        case 1: GRGL_PULL(16, 1);
        case 2: GRGL_PULL(16, 2);
        case 4: GRGL_PULL(16, 4);
        case 8: return wideKernel<T, XOR, 8>(blk, unr);
        case 16: return wideKernel<T, XOR, 16>(blk, unr);
        case 32: return wideKernel<T, XOR, 32>(blk, unr);
        default: GRGL_PULL(16, 64);
        }
    }
#undef GRGL_PULL
    throw std::logic_error("no level kernel for this access width");
}

template <typename T, bool XOR> const void* tailKernelT(uint32_t vb, uint32_t tw) {
#define GRGL_TAIL(VB, TW) return reinterpret_cast<const void*>(&pull_tail_kernel<T, VB, TW, XOR>)
    if (vb == 4) {
        if constexpr (sizeof(T) <= 4) {
            if (tw == 1) GRGL_TAIL(4, 1);
            if (tw == 2) GRGL_TAIL(4, 2);
            if (tw == 4) GRGL_TAIL(4, 4);
            GRGL_TAIL(4, 8);
        }
    } else if (vb == 8) {
        if constexpr (sizeof(T) <= 8) {
            if (tw == 1) GRGL_TAIL(8, 1);
            if (tw == 2) GRGL_TAIL(8, 2);
            if (tw == 4) GRGL_TAIL(8, 4);
            GRGL_TAIL(8, 8);
        }
    } else {
        switch (tw) {
        case 1: GRGL_TAIL(16, 1);
        case 2: GRGL_TAIL(16, 2);
        case 4: GRGL_TAIL(16, 4);
        case 8: GRGL_TAIL(16, 8);
        case 16: GRGL_TAIL(16, 16);
        case 32: GRGL_TAIL(16, 32);
        default: GRGL_TAIL(16, 64);
        }
    }
#undef GRGL_TAIL
    throw std::logic_error("no tail kernel for this access width");
}

// Grouped-level kernel for slots of 4 << cls bytes (a whole row per lane) and blocks of `blk`
// threads (256 or 1024), or null when a row of that size cannot hold one element of the type.
template <typename T, bool XOR, int BLK> const void* groupKernelB(uint32_t cls) {
#define GRGL_GROUP(RB) return reinterpret_cast<const void*>(&pull_group_kernel<T, RB, XOR, BLK>)
    switch (cls) {
    case 0:
        if constexpr (sizeof(T) <= 4) GRGL_GROUP(4);
        return nullptr;
    case 1: GRGL_GROUP(8);
    default:
        if constexpr (BLK == 256) {
            if (cls == 2) GRGL_GROUP(16);
            if (cls == 3) GRGL_GROUP(32);
            GRGL_GROUP(64);
        }
        return nullptr;
    }
#undef GRGL_GROUP
}

template <typename T, bool XOR> const void* groupKernelT(uint32_t cls, uint32_t blk) {
    if (blk == 256) return groupKernelB<T, XOR, 256>(cls);
    // 1024-thread blocks for slots of 4 and 8 bytes only: wider rows need more registers
    // than such blocks leave (and run faster in 256-thread blocks).
    if (cls > 1) return nullptr;
    return groupKernelB<T, XOR, 1024>(cls);
}

template <typename T, bool XOR> const void* groupPrefixKernelT(uint32_t cls) {
#define GRGL_PREFIX(RB) return reinterpret_cast<const void*>(&group_prefix_kernel<T, RB, XOR>)
    switch (cls) {
    case 0:
        if constexpr (sizeof(T) <= 4) GRGL_PREFIX(4);
        return nullptr;
    case 1: GRGL_PREFIX(8);
    case 2: GRGL_PREFIX(16);
    case 3: GRGL_PREFIX(32);
    default: GRGL_PREFIX(64);
    }
#undef GRGL_PREFIX
}

const void* groupPrefixKernel(TypeClass tc, uint32_t cls) {
    switch (tc) {
    case kF64: return groupPrefixKernelT<double, false>(cls);
    case kF32: return groupPrefixKernelT<float, false>(cls);
    case kW64: return groupPrefixKernelT<uint64_t, false>(cls);
    case kW32: return groupPrefixKernelT<uint32_t, false>(cls);
    case kW16: return groupPrefixKernelT<uint16_t, false>(cls);
    case kW8: return groupPrefixKernelT<uint8_t, false>(cls);
    case kX8: return groupPrefixKernelT<uint8_t, true>(cls);
    }
    throw std::logic_error("bad type class");
}

const void* groupKernel(TypeClass tc, uint32_t cls, uint32_t blk) {
    switch (tc) {
    case kF64: return groupKernelT<double, false>(cls, blk);
    case kF32: return groupKernelT<float, false>(cls, blk);
    case kW64: return groupKernelT<uint64_t, false>(cls, blk);
    case kW32: return groupKernelT<uint32_t, false>(cls, blk);
    case kW16: return groupKernelT<uint16_t, false>(cls, blk);
    case kW8: return groupKernelT<uint8_t, false>(cls, blk);
    case kX8: return groupKernelT<uint8_t, true>(cls, blk);
    }
    throw std::logic_error("bad type class");
}

const void* tailKernel(TypeClass tc, uint32_t vb, uint32_t tw) {
    switch (tc) {
    case kF64: return tailKernelT<double, false>(vb, tw);
    case kF32: return tailKernelT<float, false>(vb, tw);
    case kW64: return tailKernelT<uint64_t, false>(vb, tw);
    case kW32: return tailKernelT<uint32_t, false>(vb, tw);
    case kW16: return tailKernelT<uint16_t, false>(vb, tw);
    case kW8: return tailKernelT<uint8_t, false>(vb, tw);
    case kX8: return tailKernelT<uint8_t, true>(vb, tw);
    }
    throw std::logic_error("bad type class");
}

// Block size and loads in flight the level kernel of (type, vb, tw) runs with, for the tuned
// request (0 = default: 256 threads, 4 loads): only wide-team shapes of 4- and 8-byte types
// take other values.
void tileShape(TypeClass tc, uint32_t vb, uint32_t tw, const Tuning& t, uint32_t& blk, uint32_t& unr) {
    blk = kBlock;
    unr = 4;
    if (!tileTunable(vb, tw) || tc == kW16 || tc == kW8 || tc == kX8) return;
    if (t.blockThreads == 512) blk = 512;
    if (t.loadsInFlight == 8) unr = 8;
}

const void* pullKernel(TypeClass tc, uint32_t vb, uint32_t tw, bool h32, uint32_t blk, uint32_t unr) {
    switch (tc) {
    case kF64: return pullKernelT<double, false>(vb, tw, h32, blk, unr);
    case kF32: return pullKernelT<float, false>(vb, tw, h32, blk, unr);
    case kW64: return pullKernelT<uint64_t, false>(vb, tw, h32, blk, unr);
    case kW32: return pullKernelT<uint32_t, false>(vb, tw, h32, blk, unr);
    case kW16: return pullKernelT<uint16_t, false>(vb, tw, h32, blk, unr);
    case kW8: return pullKernelT<uint8_t, false>(vb, tw, h32, blk, unr);
    case kX8: return pullKernelT<uint8_t, true>(vb, tw, h32, blk, unr);
    }
    throw std::logic_error("bad type class");
}

// Chained kernels exist for the narrow-team shapes (TW <= 4, or accesses under 16 bytes) of
// 4- and 8-byte types; H32 for 32-byte rows.  Null: no chained kernel for this shape.
template <typename T> const void* chainKernelT(uint32_t vb, uint32_t tw, bool h32) {
#define GRGL_CHAIN(VB, TW) return reinterpret_cast<const void*>(&pull_chain_kernel<T, VB, TW, false, false>)
#define GRGL_CHAIN32(VB, TW) return reinterpret_cast<const void*>(&pull_chain_kernel<T, VB, TW, false, true>)
    if (h32) {
        if (vb == 16 && tw == 1) GRGL_CHAIN32(16, 1);
        if (vb == 16 && tw == 2) GRGL_CHAIN32(16, 2);
        if (vb == 8 && tw == 4) GRGL_CHAIN32(8, 4);
        if (vb == 8 && tw == 8) GRGL_CHAIN32(8, 8);
        if constexpr (sizeof(T) <= 4) {
            if (vb == 4 && tw == 8) GRGL_CHAIN32(4, 8);
        }
        return nullptr;
    }
    if (vb == 4) {
        if constexpr (sizeof(T) <= 4) {
            if (tw == 1) GRGL_CHAIN(4, 1);
            if (tw == 2) GRGL_CHAIN(4, 2);
            if (tw == 4) GRGL_CHAIN(4, 4);
            if (tw == 8) GRGL_CHAIN(4, 8);
        }
    } else if (vb == 8) {
        if (tw == 1) GRGL_CHAIN(8, 1);
        if (tw == 2) GRGL_CHAIN(8, 2);
        if (tw == 4) GRGL_CHAIN(8, 4);
        if (tw == 8) GRGL_CHAIN(8, 8);
    } else if (vb == 16) {
        if (tw == 1) GRGL_CHAIN(16, 1);
        if (tw == 2) GRGL_CHAIN(16, 2);
        if (tw == 4) GRGL_CHAIN(16, 4);
    }
#undef GRGL_CHAIN
#undef GRGL_CHAIN32
    return nullptr;
}

const void* chainKernel(TypeClass tc, uint32_t vb, uint32_t tw, bool h32) {
    switch (tc) {
    case kF64: return chainKernelT<double>(vb, tw, h32);
    case kF32: return chainKernelT<float>(vb, tw, h32);
    case kW64: return chainKernelT<uint64_t>(vb, tw, h32);
    case kW32: return chainKernelT<uint32_t>(vb, tw, h32);
    default: return nullptr;
    }
}

// Resident blocks of `func` on the whole device (occupancy per compute unit x units), cached.
uint32_t residentBlocks(int device, const void* func) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, uint32_t> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({device, func});
    if (it != cache.end()) return it->second;
    int perCu = 0, cus = 0;
    GRGL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, func, kBlock, 0));
    GRGL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const uint32_t n = (uint32_t)std::max(1, perCu) * (uint32_t)std::max(1, cus);
    cache[{device, func}] = n;
    return n;
}

// Every dispatch is checked on the host before it is issued: a grid of at least one block in
// each dimension, within the hardware's limits (x < 2^31, y and z < 2^16), and a block size the
// kernels are built for.  A violation is an internal error reported as GRGL_ERR_DEVICE; no
// malformed dispatch reaches the device.
void checkLaunch(dim3 grid, uint32_t block, const char* what) {
    const bool ok = grid.x >= 1 && grid.y >= 1 && grid.z == 1 && grid.x <= 0x7fffffffU && grid.y <= 65535U &&
                    (block == 256 || block == 512 || block == 1024);
    if (!ok)
        throw DeviceError(std::string("internal error: malformed ") + what + " launch (grid " + std::to_string(grid.x) +
                          " x " + std::to_string(grid.y) + " x " + std::to_string(grid.z) + ", block " +
                          std::to_string(block) + ")");
}

template <typename T, bool XOR> void launchSeedT(dim3 grid, const SeedArgs& a, bool vr, bool vm, hipStream_t s) {
    constexpr int V = 16 / (int)sizeof(T);
    if (vr && vm) hipLaunchKernelGGL((seed_rows_kernel<T, XOR, V, V>), grid, dim3(kBlock), 0, s, a);
    else if (vr) hipLaunchKernelGGL((seed_rows_kernel<T, XOR, V, 1>), grid, dim3(kBlock), 0, s, a);
    else if (vm) hipLaunchKernelGGL((seed_rows_kernel<T, XOR, 1, V>), grid, dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL((seed_rows_kernel<T, XOR, 1, 1>), grid, dim3(kBlock), 0, s, a);
}
void launchSeed(TypeClass tc, dim3 grid, const SeedArgs& a, bool vr, bool vm, hipStream_t s) {
    checkLaunch(grid, kBlock, "seed");
    switch (tc) {
    case kF64: launchSeedT<double, false>(grid, a, vr, vm, s); break;
    case kF32: launchSeedT<float, false>(grid, a, vr, vm, s); break;
    case kW64: launchSeedT<uint64_t, false>(grid, a, vr, vm, s); break;
    case kW32: launchSeedT<uint32_t, false>(grid, a, vr, vm, s); break;
    case kW16: launchSeedT<uint16_t, false>(grid, a, vr, vm, s); break;
    case kW8: launchSeedT<uint8_t, false>(grid, a, vr, vm, s); break;
    case kX8: launchSeedT<uint8_t, true>(grid, a, vr, vm, s); break;
    }
}

template <typename T, bool XOR> void launchStackedT(bool seed, dim3 grid, const StackArgs& a, bool vr, bool vm,
                                                    hipStream_t s) {
    constexpr int V = 16 / (int)sizeof(T);
    const int mode = a.per != 1 ? 0 : (a.accumulate ? 2 : 1);
#define GRGL_STACK(R, M)                                                                                            \
    do {                                                                                                            \
        if (seed) hipLaunchKernelGGL((seed_stacked_kernel<T, XOR, R, M>), grid, dim3(kBlock), 0, s, a);            \
        else if (mode == 1) hipLaunchKernelGGL((readout_stacked_kernel<T, XOR, R, M, 1>), grid, dim3(kBlock), 0, s, a); \
        else if (mode == 2) hipLaunchKernelGGL((readout_stacked_kernel<T, XOR, R, M, 2>), grid, dim3(kBlock), 0, s, a); \
        else hipLaunchKernelGGL((readout_stacked_kernel<T, XOR, R, M, 0>), grid, dim3(kBlock), 0, s, a);           \
    } while (0)
    if (vr && vm) GRGL_STACK(V, V);
    else if (vr) GRGL_STACK(V, 1);
    else if (vm) GRGL_STACK(1, V);
    else GRGL_STACK(1, 1);
#undef GRGL_STACK
}
void launchStacked(TypeClass tc, bool seed, dim3 grid, const StackArgs& a, bool vr, bool vm, hipStream_t s) {
    checkLaunch(grid, kBlock, seed ? "stacked seed" : "stacked read-out");
    switch (tc) {
    case kF64: launchStackedT<double, false>(seed, grid, a, vr, vm, s); break;
    case kF32: launchStackedT<float, false>(seed, grid, a, vr, vm, s); break;
    case kW64: launchStackedT<uint64_t, false>(seed, grid, a, vr, vm, s); break;
    case kW32: launchStackedT<uint32_t, false>(seed, grid, a, vr, vm, s); break;
    case kW16: launchStackedT<uint16_t, false>(seed, grid, a, vr, vm, s); break;
    case kW8: launchStackedT<uint8_t, false>(seed, grid, a, vr, vm, s); break;
    case kX8: launchStackedT<uint8_t, true>(seed, grid, a, vr, vm, s); break;
    }
}

template <typename T, bool XOR, int MODE> void launchReadoutM(dim3 grid, const ReadoutArgs& a, bool vr, bool vm,
                                                             hipStream_t s) {
    constexpr int V = 16 / (int)sizeof(T);
    if (vr && vm) hipLaunchKernelGGL((readout_kernel<T, XOR, V, V, MODE>), grid, dim3(kBlock), 0, s, a);
    else if (vr) hipLaunchKernelGGL((readout_kernel<T, XOR, V, 1, MODE>), grid, dim3(kBlock), 0, s, a);
    else if (vm) hipLaunchKernelGGL((readout_kernel<T, XOR, 1, V, MODE>), grid, dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL((readout_kernel<T, XOR, 1, 1, MODE>), grid, dim3(kBlock), 0, s, a);
}
template <typename T, bool XOR> void launchReadoutT(dim3 grid, const ReadoutArgs& a, bool vr, bool vm, hipStream_t s) {
    if (a.per != 1) launchReadoutM<T, XOR, 0>(grid, a, vr, vm, s);
    else if (a.accumulate) launchReadoutM<T, XOR, 2>(grid, a, vr, vm, s);
    else launchReadoutM<T, XOR, 1>(grid, a, vr, vm, s);
}
void launchReadout(TypeClass tc, dim3 grid, const ReadoutArgs& a, bool vr, bool vm, hipStream_t s) {
    checkLaunch(grid, kBlock, "read-out");
    switch (tc) {
    case kF64: launchReadoutT<double, false>(grid, a, vr, vm, s); break;
    case kF32: launchReadoutT<float, false>(grid, a, vr, vm, s); break;
    case kW64: launchReadoutT<uint64_t, false>(grid, a, vr, vm, s); break;
    case kW32: launchReadoutT<uint32_t, false>(grid, a, vr, vm, s); break;
    case kW16: launchReadoutT<uint16_t, false>(grid, a, vr, vm, s); break;
    case kW8: launchReadoutT<uint8_t, false>(grid, a, vr, vm, s); break;
    case kX8: launchReadoutT<uint8_t, true>(grid, a, vr, vm, s); break;
    }
}

void launchInit(TypeClass tc, const InitArgs& a, hipStream_t s) {
    const uint64_t n = (uint64_t)(a.rowEnd - a.rowBegin) * a.K;
    if (n == 0) return;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + kBlock - 1) / kBlock, 65536);
    checkLaunch(dim3(blocks), kBlock, "row initialisation");
    switch (tc) {
    case kF64: hipLaunchKernelGGL((init_rows_kernel<double>), dim3(blocks), dim3(kBlock), 0, s, a); break;
    case kF32: hipLaunchKernelGGL((init_rows_kernel<float>), dim3(blocks), dim3(kBlock), 0, s, a); break;
    case kW64: hipLaunchKernelGGL((init_rows_kernel<uint64_t>), dim3(blocks), dim3(kBlock), 0, s, a); break;
    case kW32: hipLaunchKernelGGL((init_rows_kernel<uint32_t>), dim3(blocks), dim3(kBlock), 0, s, a); break;
    case kW16: hipLaunchKernelGGL((init_rows_kernel<uint16_t>), dim3(blocks), dim3(kBlock), 0, s, a); break;
    default: hipLaunchKernelGGL((init_rows_kernel<uint8_t>), dim3(blocks), dim3(kBlock), 0, s, a); break;
    }
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15U) == 0; }

// Tile height: TK = min(64, pow2ceil(K)); TC = tile elements / TK.
uint32_t tileShift(uint32_t K) { return std::min<uint32_t>(6, log2u(pow2ceil(K))); }

// Elements of the transpose tiles (kTileElems<T> / kSeedTileElems<T> of the kernels).
uint32_t tileElems(size_t es) { return es == 1 ? 4096U : 8192U / (uint32_t)es; }
uint32_t seedTileElems(size_t es) { return es == 8 ? 2048U : 4096U; }

// Row-side 16-byte accesses need a 16-byte aligned row stride and request offset, and a
// request at least 16 bytes wide: only such requests own whole 16-byte slots (stackOffsets
// rounds their slots up), so a vector running past K touches padding, never another
// request's columns.
bool rowVectorOk(uint64_t ld, uint32_t kOffset, uint32_t K, size_t es) {
    return ld % 16 == 0 && ((uint64_t)kOffset * es) % 16 == 0 && (uint64_t)K * es >= 16;
}

void seedRows(TypeClass tc, size_t es, const void* in, uint64_t inCols, uint32_t K, char* rows, uint64_t ld,
              uint32_t rowBase, uint32_t kOffset, uint32_t D, uint32_t per, const uint32_t* colMap, bool accumulate,
              hipStream_t s) {
    if (D == 0 || K == 0) return;
    SeedArgs a;
    a.in = in;
    a.inCols = inCols;
    a.K = K;
    a.rows = rows;
    a.ld = ld;
    a.rowBase = rowBase;
    a.kOffset = kOffset;
    a.D = D;
    a.per = per;
    a.colMap = colMap;
    a.accumulate = accumulate ? 1 : 0;
    a.tkShift = tileShift(K);
    a.normalizeBool = tc == kX8 ? 1 : 0;
    // The kernels' fixed access counts assume a tile line of at least one row-side vector.
    const bool vr = rowVectorOk(ld, kOffset, K, es) && (1ULL << a.tkShift) * es >= 16;
    const bool vm = colMap == nullptr && per == 1 && (inCols * es) % 16 == 0 && aligned16(in);
    const uint32_t TK = 1U << a.tkShift, TC = seedTileElems(es) / TK;
    dim3 grid((D + TC - 1) / TC, (K + TK - 1) / TK);
    launchSeed(tc, grid, a, vr, vm, s);
}

void readoutRows(TypeClass tc, size_t es, const char* rows, uint64_t ld, uint32_t kOffset, uint32_t K,
                 const uint32_t* map, uint32_t per, uint64_t C, void* out, uint64_t outCols, bool accumulate,
                 hipStream_t s) {
    if (C == 0 || K == 0) return;
    ReadoutArgs a;
    a.rows = rows;
    a.ld = ld;
    a.kOffset = kOffset;
    a.K = K;
    a.map = map;
    a.per = per;
    a.C = C;
    a.out = out;
    a.outCols = outCols;
    a.accumulate = accumulate ? 1 : 0;
    a.tkShift = tileShift(K);
    const bool vr = rowVectorOk(ld, kOffset, K, es) && (1ULL << a.tkShift) * es >= 16;
    const bool vm = (outCols * es) % 16 == 0 && aligned16(out);
    const uint32_t TK = 1U << a.tkShift, TC = tileElems(es) / TK;
    dim3 grid((uint32_t)((C + TC - 1) / TC), (K + TK - 1) / TK);
    launchReadout(tc, grid, a, vr, vm, s);
}

// Stacked transposes (one launch for all requests of a traversal), see kernels.h.  `ptrs`
// are the request matrices (K x cols each), `seed` selects the direction of the transpose.
// Requests must be sorted by kOffset and fit kMaxStack.
void stackedTranspose(TypeClass tc, size_t es, bool seed, const std::vector<TravRequest>& reqs, bool useOut,
                      const RowGeometry& geo, char* rows, uint32_t rowBase, uint64_t C, uint64_t cols, uint32_t per,
                      const uint32_t* map, bool accumulate, hipStream_t s) {
    if (C == 0) return;
    StackArgs a{};
    a.rows = rows;
    a.ld = geo.ld;
    a.rowBase = rowBase;
    a.kCols = (uint32_t)(geo.rowBytes / es);
    a.C = C;
    a.cols = cols;
    a.per = per;
    a.map = map;
    a.tkShift = tileShift(a.kCols);
    a.accumulate = accumulate ? 1 : 0;
    a.normalizeBool = (seed && tc == kX8) ? 1 : 0;
    a.numReq = (uint32_t)reqs.size();
    bool vm = per == 1 || !seed;
    for (size_t i = 0; i < reqs.size(); i++) {
        const void* p = useOut ? reqs[i].out : reqs[i].in;
        a.req[i].ptr = p;
        a.req[i].kOffset = reqs[i].kOffset;
        a.req[i].K = reqs[i].K;
        vm = vm && aligned16(p);
    }
    vm = vm && (cols * es) % 16 == 0;
    if (seed) vm = vm && per == 1;
    const bool vr = geo.ld % 16 == 0 && geo.rowBytes % 16 == 0 && (1ULL << a.tkShift) * es >= 16;
    const uint32_t TK = 1U << a.tkShift, TC = (seed ? seedTileElems(es) : tileElems(es)) / TK;
    dim3 grid((uint32_t)((C + TC - 1) / TC), (a.kCols + TK - 1) / TK);
    launchStacked(tc, seed, grid, a, vr, vm, s);
}

// Requests sorted by kOffset, or empty when they cannot be stacked.
std::vector<TravRequest> stackable(const std::vector<TravRequest>& reqs) {
    if (reqs.size() < 2 || reqs.size() > (size_t)kMaxStack) return {};
    std::vector<TravRequest> sorted(reqs);
    std::sort(sorted.begin(), sorted.end(),
              [](const TravRequest& x, const TravRequest& y) { return x.kOffset < y.kOffset; });
    for (size_t i = 1; i < sorted.size(); i++)
        if (sorted[i].kOffset < sorted[i - 1].kOffset + sorted[i - 1].K) return {}; // overlapping columns
    return sorted;
}

// Grouped levels replace the per-level launches of their levels unless a traversal asks for
// one launch per level (chain = 1) or a chained launch (chain = 2).
constexpr bool kGroupByDefault = true;

// Rows of up to this many bytes take grouped levels by default; wider rows when the tuning
// asks for them.  UP: 512 (FP32 K <= 128); on the 200k x 1M graph 64-byte slices save 7-9%
// at K = 17-32 and 3-6% at K = 48-128.  DOWN: 256 (FP32 K <= 64), rows over 64 bytes over its
// small levels only (DeviceDirPlan::lateWide): 2-4% saved at K = 17-64 on the 200k x 1M graph,
// 0.5-3% on the constructed 100k x 500k graph, none at K = 128-256 (over all its levels the
// slices lost 1-8% at K >= 48).
uint32_t groupDefaultMaxRowBytes(const DeviceDirPlan& plan) { return plan.down ? 256 : 512; }

Tuning chooseTuning(grgl_graph* g, grgl_direction dir, TypeClass tc, uint32_t K, const RowGeometry& geo) {
    Tuning t;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        auto it = g->tuning.find(std::make_tuple((int)dir, (int)tc, K));
        if (it != g->tuning.end()) t = it->second;
    }
    uint32_t tw = t.teamLanes;
    // Default: 16 lanes (256-byte row slices) at most.  Blocks are dispatched K-tile by
    // K-tile, so narrower slices keep each tile's gathered rows cache resident.  Rows of
    // one or two slices take 8 lanes (4 on rows of at most 8 bytes): narrower accesses,
    // then idle lanes, put fewer rows in each wavefront, so a level's rows reach more
    // compute units and more of them are in flight at once (latency-bound small K: on the
    // 200k x 1M graph FP32 DOWN K = 1 goes from 0.35 to 0.22 ms, K = 8 from 0.36 to 0.25 ms).
    // Rows of 256 B - 1 KB (FP32 K = 64..256) take 8 lanes UP and 16 DOWN.  Over seven graphs
    // (tree-direct 20k x 100k with 1,000 trees, 100k x 500k, 200k x 1M; haplotype-clustered
    // 20k x 100k, 40k x 200k, 100k x 500k, 20k x 100k from 1,000 trees) that keeps every
    // K = 64..256 within 10% of the best team width.  16 lanes UP lost up to 45% (K = 64 on
    // the 1,000-tree graph, 0.66 vs 0.46 ms); 8 lanes DOWN lost up to 17% (K = 256 on the
    // clustered 100k x 500k graph, 2.20 vs 1.87 ms).
    // 16-byte rows (FP32 K = 4) take 4 lanes UP: 8 lanes lost up to 11% on the same graphs,
    // 4 lanes at most 5%; DOWN keeps 8 (4 lanes lost 11% on the 200k x 1M graph).
    if (tw == 0) {
        if (geo.lanes < 3)
            tw = (geo.rowBytes <= 8 || (dir == GRGL_DIRECTION_UP && geo.rowBytes <= 16)) ? 4 : 8;
        else if (geo.rowBytes >= 256 && geo.rowBytes <= 1024)
            tw = dir == GRGL_DIRECTION_UP ? 8 : 16;
        else
            tw = std::min<uint32_t>(16, pow2ceil(geo.lanes));
    }
    t.teamLanes = std::min<uint32_t>(maxTeamLanes(geo), pow2ceil(tw));
    // Grouped levels (where the direction has them for this row size, up to its default width),
    // else one launch per level.
    if (t.chain == 0)
        t.chain = kGroupByDefault && geo.rowBytes <= groupDefaultMaxRowBytes(dir == GRGL_DIRECTION_UP ? g->up : g->down)
                      ? 3
                      : 1;
    // Rows of 128 B - 1 KB (FP32 K = 32..256): 8 row loads in flight per batch; the autotuner
    // picks it for these on the 200k x 1M graph (DOWN K = 32: 0.47 -> 0.41 ms, K = 64: 0.75 ->
    // 0.70 ms).  Wider rows keep 4 (the 511-RHS traversal measures no gain).
    if (t.loadsInFlight == 0 && geo.rowBytes >= 128 && geo.rowBytes <= 1024) t.loadsInFlight = 8;
    return t;
}

// First level of the trailing run that one block handles (numLevels: none).  A level
// qualifies when its light rows take at most two passes of the block's teams, it has at
// most two heavy rows, and its rows and gathered rows come to at most 64 KB per column
// slice set; the run needs at least two levels to save a launch.
uint32_t tailStart(const DeviceDirPlan& plan, const LaunchTable& tab, const RowGeometry& geo, uint32_t tw,
                   bool orderPreserving) {
    const uint32_t teams = kBlock / tw;
    uint32_t l = plan.numLevels;
    while (l > plan.firstPulledLevel && plan.numLevels - l < (uint32_t)kMaxTailLevels) {
        const uint32_t k = l - 1;
        const uint64_t nodes = plan.levelStart[k + 1] - plan.levelStart[k];
        const uint64_t heavy = orderPreserving ? 0 : tab.heavyCount[k];
        const uint64_t bytes = (tab.levelEntries[k] + nodes) * (uint64_t)geo.rowBytes;
        if (nodes - heavy > 2ULL * teams || heavy > 2 || bytes > (64ULL << 10)) break;
        l = k;
    }
    return plan.numLevels - l >= 2 ? l : plan.numLevels;
}

// One level-kernel launch: kernel address, grid and arguments (a PullArgs, or the TailArgs of
// the single-block tail launch), plus the algorithmic bytes it moves (kernel timing).
struct LevelLaunch {
    const void* func = nullptr;
    dim3 grid;
    uint32_t block = kBlock;
    uint32_t shmem = 0;  // dynamic LDS bytes
    bool tail = false;
    bool group = false;  // a grouped-levels launch or its prefix launch (GroupArgs)
    bool prefix = false; // the prefix launch
    PullArgs pull{};
    TailArgs tailArgs{};
    GroupArgs groupArgs{};
    double bytes = 0.0;
    void* argPtr() { return tail ? (void*)&tailArgs : (group ? (void*)&groupArgs : (void*)&pull); }
    const char* what() const {
        return tail ? "tail level" : (group ? (prefix ? "grouped prefix" : "grouped levels") : "level");
    }
};


// Threads per block of a grouped launch: block_threads 256 or 1024 when the tuning names one,
// else 1024 for slots of 4 and 8 bytes (the largest groups, whose rows then spread over more
// lanes in phase 0 and per level); 256 for wider slots, which have no 1024-thread form.
uint32_t groupBlockThreads(const Tuning& t, TypeClass tc, uint32_t cls) {
    const uint32_t want = (t.blockThreads == 256 || t.blockThreads == 1024) ? t.blockThreads : (cls <= 1 ? 1024 : 256);
    return (want == 1024 && groupKernel(tc, cls, 1024)) ? 1024 : 256;
}

// The level-sequence form a call with tuning `t` takes (grouped levels fall back to one
// launch per level where the direction has none for this row size), and for grouped levels
// the grouped launch's block size.
void reportChain(const DeviceDirPlan& plan, const LaunchTable& tab, TypeClass tc, const RowGeometry& geo,
                 const Tuning& t, grgl_tuning& out);

// Slot class of a row of `rowBytes`: slots of 4 << class bytes (64 for wider rows, which
// run as 64-byte column slices).
uint32_t groupClassOf(uint32_t rowBytes) {
    uint32_t c = 0;
    while ((4U << c) < rowBytes && c + 1 < kDeviceGroupClasses) c++;
    return c;
}

// Column slices of a grouped launch: one for rows of at most 64 bytes, else one per 64 bytes.
uint32_t groupSlices(const RowGeometry& geo) { return geo.rowBytes <= 64 ? 1 : (geo.rowBytes + 63) / 64; }

// The grouped-launch plan a traversal takes, or null: the launch table's heavy threshold
// equal to the plan's, not order-preserving (whose lists are all light), no per-launch
// timing, and rows of at most groupDefaultMaxRowBytes unless the tuning asks for grouped
// levels (chain 3) whatever the width; wider rows run as 64-byte slices, each inside the row.
const DeviceGroupPlan* groupFor(const DeviceDirPlan& plan, const LaunchTable& tab, TypeClass tc,
                                const RowGeometry& geo, const Tuning& t, bool orderPreserving, bool timing) {
    const bool want =
        t.chain == 3 || (t.chain == 0 && kGroupByDefault && geo.rowBytes <= groupDefaultMaxRowBytes(plan));
    if (!want || timing || orderPreserving) return nullptr;
    if (geo.rowBytes > 64 && (uint64_t)groupSlices(geo) * 64 > geo.ld) return nullptr;
    const uint32_t cls = groupClassOf(geo.rowBytes);
    // DOWN rows over 64 bytes: only the small levels (the larger ones move wide rows faster
    // one launch each).
    const DeviceGroupPlan& gp = plan.down && geo.rowBytes > 64 ? plan.lateWide : plan.groups[cls];
    if (gp.empty() || gp.heavyFanin != tab.heavyFanin || !groupKernel(tc, cls, 256)) return nullptr;
    return &gp;
}

void reportChain(const DeviceDirPlan& plan, const LaunchTable& tab, TypeClass tc, const RowGeometry& geo,
                 const Tuning& t, grgl_tuning& out) {
    out.chain = t.chain;
    if (t.chain == 3) {
        if (groupFor(plan, tab, tc, geo, t, false, false))
            out.block_threads = groupBlockThreads(t, tc, groupClassOf(geo.rowBytes));
        else
            out.chain = 1;
    }
}

// Output fusion of a traversal (PullArgs::out): the last level's rows written straight to
// their output elements.
struct OutFusion {
    char* out;
    const uint32_t* outCol;
    uint32_t outCols;
    uint32_t K;
};

// Seed fold of a one-column traversal (PullArgs::readRows, copy*): the first pulled level reads
// its seed rows (all its lists read) from the input, and its launch copies the input to the
// seed rows [seedBegin, seedBegin + bytes / ld) for the later levels, so there is no seed launch.
struct SeedFold {
    const char* in;
    uint32_t seedBegin;
    uint64_t bytes;
};

// The launches of every pulled level of one direction, in order (the trailing small levels as
// one tail launch).  With `timing`, every level gets a launch of its own and its bytes.  With
// `fuse`, the last level's level-kernel launches write its rows to the output when a row is one
// column slice (*fused is then set; a last level in the tail launch is not fused).  With `fold`,
// the first pulled level's level-kernel launches read the seeds from the input and the first
// of them copies them (*folded is then set; a first level in a grouped or tail launch is not).
std::vector<LevelLaunch> levelLaunches(const DeviceDirPlan& plan, const LaunchTable& tab, TypeClass tc,
                                       const RowGeometry& geo, const Tuning& tuning, bool readSelf,
                                       bool orderPreserving, uint32_t K, size_t elemSize, char* rows, bool timing,
                                       size_t groupScratch, const OutFusion* fuse = nullptr, bool* fused = nullptr,
                                       const SeedFold* fold = nullptr, bool* folded = nullptr) {
    std::vector<LevelLaunch> out;
    bool copyAttached = false;
    const uint32_t tw = tuning.teamLanes;
    const bool xcdMap = tuning.xcdMap == 1;
    const uint32_t tiles = levelTiles(geo, tw, elemSize);
    const uint32_t vb = levelVb(geo, tw, elemSize);
    uint32_t blk = kBlock, unr = 4;
    tileShape(tc, vb, tw, tuning, blk, unr);
    const uint32_t teamsPerBlock = blk / tw;
    // The single-block tail gains nothing from spreading rows over idle lanes: rows of one
    // access take single-lane teams there (256 rows per pass instead of 256 / tw).
    const uint32_t tailTw = geo.lanes == 1 ? 1 : tw;
    const DeviceGroupPlan* gp = groupFor(plan, tab, tc, geo, tuning, orderPreserving, timing);
    const uint32_t tailFrom = (timing || gp) ? plan.numLevels : tailStart(plan, tab, geo, tailTw, orderPreserving);
    for (uint32_t l = plan.firstPulledLevel; l < tailFrom; l++) {
      if (gp && l >= gp->levelBegin && l < gp->levelEnd) {
        if (l != gp->levelBegin) continue;
        LevelLaunch L;
        L.group = true;
        L.block = groupBlockThreads(tuning, tc, groupClassOf(geo.rowBytes));
        L.shmem = gp->ldsBytes;
        GroupArgs& a = L.groupArgs;
        a.rows = rows;
        a.ld = geo.ld;
        a.blockRow = gp->blockRow;
        a.blockSeg = gp->blockSeg;
        a.blockHeavy = gp->blockHeavy;
        a.blockImport = gp->blockImport;
        a.segRow = gp->segRow;
        a.segHeavy = gp->segHeavy;
        a.segHeavyOrd = gp->segHeavyOrd;
        a.row = gp->row;
        a.off = gp->off;
        a.pre = gp->pre;
        a.src = gp->src;
        a.remOff = gp->remOff;
        a.rem = gp->rem;
        a.heavyRow = gp->heavyRow;
        a.importRow = gp->importRow;
        a.heavyFanin = gp->heavyFanin;
        a.readSelf = readSelf ? 1 : 0;
        a.scratch = rows + groupScratch;
        a.numRows = gp->numRows;
        a.numHeavy = gp->numHeavy;
        a.lightBlocks = (gp->numRows + kPrefixRowsPerBlock - 1) / kPrefixRowsPerBlock;
        // Rows of at most 8 bytes with long prefixes (4 entries a row or more on average: UP,
        // whose grouped rows read many samples first): the prefix sums over the whole device
        // first, so that no block sums long lists a few rows per thread; otherwise (DOWN, whose
        // prefixes are a row's few mutation seeds, and wider rows) in the grouped launch.
        // UP K = 1 on the 200k x 1M graph: 0.130 -> 0.124 ms; DOWN K = 2 on the constructed
        // 100k x 500k graph: 0.153 -> 0.166 ms with a prefix launch.
        a.prefixDone = (geo.rowBytes <= 8 && gp->prefixEntries >= 4ULL * gp->numRows) ? 1 : 0;
        if (a.prefixDone) {
            LevelLaunch P = L;
            P.prefix = true;
            P.grid = dim3(a.lightBlocks + (gp->numHeavy + kBlock / 64 - 1) / (kBlock / 64));
            P.block = kBlock;
            P.shmem = 0;
            P.func = groupPrefixKernel(tc, groupClassOf(geo.rowBytes));
            out.push_back(P);
        }
        L.grid = dim3(gp->blocks, groupSlices(geo));
        L.func = groupKernel(tc, groupClassOf(geo.rowBytes), L.block);
        out.push_back(L);
        continue;
      }
      for (uint32_t ci = tab.levelChunk[l]; ci < tab.levelChunk[l + 1]; ci++) {
        const LevelChunk& ck = tab.chunks[ci];
        const uint32_t begin = ck.rowBegin, end = ck.rowEnd;
        if (end == begin) continue;
        const uint32_t heavy = orderPreserving ? 0 : ck.heavy;
        const uint32_t light = end - begin - heavy;
        LevelLaunch L;
        PullArgs& a = L.pull;
        a.rows = rows;
        a.readRows = rows;
        a.ld = geo.ld;
        a.off = plan.off;
        a.src = plan.src + ck.entryBegin;
        a.entryBase = (uint32_t)ck.entryBegin;
        a.levelBegin = begin;
        a.heavyCount = heavy;
        // Short rows: a wavefront per heavy row, 4 per block (see PullArgs::heavyPerWave),
        // for levels with many heavy rows (at least 512: a few hundred rows of thousands of
        // entries, the DOWN levels near the roots, go faster with a whole block each).
        // 32-byte rows go through the H32 kernel variant, and only where heavy rows hold at
        // least a quarter of the level's gather entries (its extra registers slow the light
        // rows of the launch).
        const bool manyHeavy = heavy >= 512;
        const bool h32 = geo.rowBytes == 32 && tw <= 8 && manyHeavy && tab.heavyEntries[l] * 4 >= tab.levelEntries[l];
        a.heavyPerWave = ((geo.rowBytes <= 16 && manyHeavy) || h32) ? 1 : 0;
        a.heavyBlocks = a.heavyPerWave ? (heavy + blk / 64 - 1) / (blk / 64) : heavy;
        a.levelEnd = end;
        a.rowBytes = geo.rowBytes;
        a.readSelf = readSelf ? 1 : 0;
        if (fuse && l + 1 == plan.numLevels && tiles == 1) {
            a.out = fuse->out;
            a.outCol = fuse->outCol;
            a.outCols = fuse->outCols;
            a.outK = fuse->K;
            if (fused) *fused = true;
        }
        const uint32_t bxCount = a.heavyBlocks + (light + teamsPerBlock - 1) / teamsPerBlock;
        a.xcdTiles = (xcdMap && tiles % 8 == 0) ? tiles : 0;
        L.grid = a.xcdTiles ? dim3(bxCount * tiles, 1) : dim3(bxCount, tiles);
        if (fold && l == plan.firstPulledLevel && tiles == 1 && geo.ld == elemSize) {
            // Seed row r of the input sits at in + (r - seedBegin) * ld: read it there.
            a.readRows = reinterpret_cast<const char*>(reinterpret_cast<uintptr_t>(fold->in) -
                                                       (uintptr_t)fold->seedBegin * geo.ld);
            if (!copyAttached) { // the level's first launch copies the seeds
                copyAttached = true;
                a.copySrc = fold->in;
                a.copyDst = rows + (uint64_t)fold->seedBegin * geo.ld;
                a.copyBytes = fold->bytes;
                const uint64_t words = fold->bytes / 4, perBlock = (uint64_t)blk * 16;
                a.copyBlocks = (uint32_t)std::min<uint64_t>(1024, (words + perBlock - 1) / perBlock);
                L.grid.x = a.copyBlocks + bxCount;
            }
            if (folded) *folded = true;
        }
        L.func = pullKernel(tc, vb, tw, h32, blk, unr);
        L.block = blk;
        if (timing) {
            const uint64_t nodes = end - begin;
            uint64_t ent = 0;
            for (uint32_t r = begin; r < end; r++) ent += plan.fanIn[r];
            const double rowB = (double)K * (double)elemSize;
            L.bytes = 4.0 * (double)(ent + nodes + 1) + rowB * (double)(ent + nodes + (readSelf ? nodes : 0));
        }
        out.push_back(L);
      }
    }
    if (tailFrom < plan.numLevels) {
        LevelLaunch L;
        L.tail = true;
        L.grid = dim3(1);
        TailArgs& t = L.tailArgs;
        t.rows = rows;
        t.ld = geo.ld;
        // Tail levels are small (at most 64 KB of rows each), so one base covers them all;
        // the tail kernel takes whole levels.
        const uint64_t tailEntry = tab.levelEntryBegin[tailFrom];
        t.off = plan.off;
        t.src = plan.src + tailEntry;
        t.entryBase = (uint32_t)tailEntry;
        t.rowBytes = geo.rowBytes;
        t.readSelf = readSelf ? 1 : 0;
        t.tiles = levelTiles(geo, tailTw, elemSize);
        t.numLevels = plan.numLevels - tailFrom; // <= kMaxTailLevels (tailStart)
        for (uint32_t i = 0; i <= t.numLevels; i++) t.levelStart[i] = plan.levelStart[tailFrom + i];
        for (uint32_t i = 0; i < t.numLevels; i++) t.heavyCount[i] = orderPreserving ? 0 : tab.heavyCount[tailFrom + i];
        L.func = tailKernel(tc, levelVb(geo, tailTw, elemSize), tailTw);
        out.push_back(L);
    }
    return out;
}

// Launches on `stream`; with timing, each launch is bracketed by events kept on the graph.
void launchDirect(grgl_graph* g, std::vector<LevelLaunch>& launches, hipStream_t stream, bool timing) {
    for (LevelLaunch& L : launches) {
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        if (timing) {
            GRGL_HIP_CHECK(hipEventCreate(&ev0));
            GRGL_HIP_CHECK(hipEventCreate(&ev1));
            GRGL_HIP_CHECK(hipEventRecord(ev0, stream));
        }
        checkLaunch(L.grid, L.block, L.what());
        void* args[1] = {L.argPtr()};
        GRGL_HIP_CHECK(hipLaunchKernel(L.func, L.grid, dim3(L.block), args, L.shmem, stream));
        if (timing) {
            GRGL_HIP_CHECK(hipEventRecord(ev1, stream));
            std::lock_guard<std::mutex> lk(g->mu);
            g->timingEvents.push_back({ev0, ev1});
            g->timingBytes.push_back(L.bytes);
        }
    }
}

// The launches as a chain of kernel nodes of a new hipGraph, instantiated.  Built node by node
// (no stream capture), so what other threads do on the device meanwhile cannot affect it.
void buildLevelGraph(std::vector<LevelLaunch>& launches, CapturedLevels& out) {
    GRGL_HIP_CHECK(hipGraphCreate(&out.graph, 0));
    hipGraphNode_t prev = nullptr;
    for (LevelLaunch& L : launches) {
        checkLaunch(L.grid, L.block, L.what());
        hipKernelNodeParams p{};
        void* args[1] = {L.argPtr()};
        p.func = const_cast<void*>(L.func);
        p.gridDim = L.grid;
        p.blockDim = dim3(L.block);
        p.sharedMemBytes = L.shmem;
        p.kernelParams = args;
        p.extra = nullptr;
        hipGraphNode_t node = nullptr;
        GRGL_HIP_CHECK(hipGraphAddKernelNode(&node, out.graph, prev ? &prev : nullptr, prev ? 1 : 0, &p));
        prev = node;
    }
    GRGL_HIP_CHECK(hipGraphInstantiate(&out.exec, out.graph, nullptr, nullptr, 0));
}

// The level sequence as ONE chained launch (pull_chain_kernel), when the shape allows it: rows
// of at most 64 bytes in one column slice on narrow teams, at most kMaxChainLevels pulled
// levels each launched whole, fewer than 2^32 gather entries, a 4- or 8-byte type, and zeroed
// chain state from the caller's buffer.  Returns false (nothing launched) otherwise.
bool launchChained(grgl_graph* g, const DeviceDirPlan& plan, const LaunchTable& tab, TypeClass tc,
                   const RowGeometry& geo, uint32_t tw, bool readSelf, bool orderPreserving, size_t elemSize,
                   char* rows, uint32_t* chainState, hipStream_t stream) {
    if (!chainState || !g->chainCtlDev || geo.rowBytes > 64) return false;
    const uint32_t vb = levelVb(geo, tw, elemSize);
    if (levelTiles(geo, tw, elemSize) != 1 || !(tw <= 4 || vb < 16)) return false;
    const uint32_t first = plan.firstPulledLevel;
    const uint32_t n = plan.numLevels > first ? plan.numLevels - first : 0;
    if (n == 0 || n > (uint32_t)kMaxChainLevels || plan.numEntries >= (1ULL << 32)) return false;
    for (uint32_t l = first; l < plan.numLevels; l++)
        if (tab.levelChunk[l + 1] - tab.levelChunk[l] != 1) return false;
    const void* func = chainKernel(tc, vb, tw, geo.rowBytes == 32);
    if (!func) return false;
    ChainArgs c{};
    c.rows = rows;
    c.ld = geo.ld;
    c.off = plan.off;
    c.src = plan.src;
    c.rowBytes = geo.rowBytes;
    c.readSelf = readSelf ? 1 : 0;
    c.numLevels = n;
    c.state = chainState;
    c.gaveUp = &g->chainCtlDev->gaveUp;
    c.abortWord = &g->chainCtlDev->abortWord;
    c.waitTicks = (uint64_t)g->chainWaitUs.load() * g->wallClockKHz / 1000;
    const uint32_t teamsPerBlock = kBlock / tw;
    uint32_t task = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t l = first + i;
        const LevelChunk& ck = tab.chunks[tab.levelChunk[l]];
        const uint32_t heavy = orderPreserving ? 0 : ck.heavy;
        const uint32_t light = ck.rowEnd - ck.rowBegin - heavy;
        // One heavy row per wavefront where the level kernels use it (levelLaunches).
        const bool manyHeavy = heavy >= 512;
        const bool h32 = geo.rowBytes == 32 && tw <= 8 && manyHeavy && tab.heavyEntries[l] * 4 >= tab.levelEntries[l];
        const bool perWave = (geo.rowBytes <= 16 && manyHeavy) || h32;
        const uint32_t heavyBlocks = perWave ? (heavy + kBlock / 64 - 1) / (kBlock / 64) : heavy;
        const uint32_t tasks = heavyBlocks + (light + teamsPerBlock - 1) / teamsPerBlock;
        if (tasks == 0) return false; // every level must hold work (a complete level implies the ones below)
        c.taskStart[i] = task;
        c.levelBegin[i] = ck.rowBegin;
        c.heavyCount[i] = heavy;
        c.heavyBlocks[i] = heavyBlocks | (perWave ? 0x80000000U : 0U);
        task += tasks;
    }
    c.taskStart[n] = task;
    c.levelBegin[n] = plan.levelStart[plan.numLevels];
    c.totalTasks = task;
    const uint32_t grid = std::min<uint32_t>(task, residentBlocks(g->device, func));
    checkLaunch(dim3(grid), kBlock, "chained level");
    void* args[1] = {&c};
    GRGL_HIP_CHECK(hipLaunchKernel(func, dim3(grid), dim3(kBlock), args, 0, stream));
    return true;
}

// Level sequence, either launched directly or replayed from a hipGraph of it.  A graph bakes
// in the row buffer address, the row geometry, the team width and the launch table, so the
// cache key holds all of them; different workspaces get different graphs.
// Returns whether the last level's rows went to the output (`fuse`: direct launches only).
// `seeds`, when given, launches the seed rows: it runs before the levels unless `fold` applies
// (direct launches whose first pulled level is a level-kernel launch), which replaces it.
bool runLevels(grgl_graph* g, grgl_direction dir, TypeClass tc, const RowGeometry& geo, const Tuning& t,
               bool readSelf, bool orderPreserving, uint32_t K, size_t es, char* rows, hipStream_t stream,
               bool launchOnly = false, uint32_t* chainState = nullptr, const OutFusion* fuse = nullptr,
               const SeedFold* fold = nullptr, const std::function<void()>& seeds = {}) {
    const DeviceDirPlan& plan = dir == GRGL_DIRECTION_UP ? g->up : g->down;
    const std::shared_ptr<const LaunchTable> tab = launchTable(g, dir);
    const bool timing = g->timing && !launchOnly;
    if (t.chain == 2 && !timing && !launchOnly) {
        if (seeds) seeds();
        if (launchChained(g, plan, *tab, tc, geo, t.teamLanes, readSelf, orderPreserving, es, rows, chainState,
                          stream))
            return false;
    }
    bool direct = !t.useGraph || timing || launchOnly;
    if (!direct && stream != nullptr) {
        // A stream the caller is capturing gets direct launches (they become part of the
        // caller's graph, which then does not depend on this library's cache).
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GRGL_HIP_CHECK(hipStreamIsCapturing(stream, &cs));
        direct = cs != hipStreamCaptureStatusNone;
    }
    const bool seeded = t.chain == 2 && !timing && !launchOnly;
    if (direct) {
        bool fused = false, folded = false;
        std::vector<LevelLaunch> launches =
            levelLaunches(plan, *tab, tc, geo, t, readSelf, orderPreserving, K, es, rows, timing,
                          groupScratchOffset(g, dir, geo), fuse, &fused, seeded ? nullptr : fold, &folded);
        if (!folded && !seeded && seeds) seeds();
        launchDirect(g, launches, stream, timing);
        return fused;
    }
    if (!seeded && seeds) seeds();
    const auto key = std::make_tuple((int)dir, (int)tc, geo.vb, t.teamLanes, geo.ld, geo.rowBytes,
                                     (int)readSelf * 2 + (int)orderPreserving + 4 * (int)(t.xcdMap == 1) +
                                         8 * (int)(t.blockThreads == 512) + 16 * (int)(t.loadsInFlight == 8) +
                                         32 * (int)tab->heavyFanin,
                                     (uintptr_t)rows, tab->chunkEntries, t.chain | (t.blockThreads << 8));
    std::shared_ptr<CapturedLevels> cap;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        auto it = g->graphs.find(key);
        if (it != g->graphs.end()) cap = it->second;
    }
    if (!cap) {
        auto fresh = std::make_shared<CapturedLevels>();
        GRGL_HIP_CHECK(hipEventCreateWithFlags(&fresh->lastLaunch, hipEventDisableTiming));
        std::vector<LevelLaunch> launches =
            levelLaunches(plan, *tab, tc, geo, t, readSelf, orderPreserving, K, es, rows, false,
                          groupScratchOffset(g, dir, geo));
        buildLevelGraph(launches, *fresh);
        std::lock_guard<std::mutex> lk(g->mu);
        if (g->graphs.size() >= kMaxCapturedGraphs) {
            // Bounded cache: drop the entries that no traversal holds and whose last launch
            // has completed (a non-blocking event query; nothing waits here).
            for (auto it = g->graphs.begin(); it != g->graphs.end();) {
                const bool idle = it->second.use_count() == 1 && hipEventQuery(it->second->lastLaunch) == hipSuccess;
                it = idle ? g->graphs.erase(it) : std::next(it);
            }
        }
        // Two threads may build the same key: the first insertion wins, and the other
        // thread's graph (never launched) is destroyed when `fresh` goes out of scope.
        cap = g->graphs.try_emplace(key, fresh).first->second;
    }
    GRGL_HIP_CHECK(hipGraphLaunch(cap->exec, stream));
    GRGL_HIP_CHECK(hipEventRecord(cap->lastLaunch, stream));
    return false;
}

} // namespace

void checkChainHealth(const grgl_graph* g) {
    if (g->chainCtl && __atomic_load_n(&g->chainCtl->gaveUp, __ATOMIC_ACQUIRE) != 0)
        throw DeviceError("a chained traversal on this graph gave up waiting (abort word set, or a wait longer than " +
                          std::to_string(g->chainWaitUs.load()) + " us): its outputs are undefined; "
                          "grgl_hip_chain_control(graph, ..., reset = 1) clears this state");
}

void enqueueTraversal(grgl_graph* g, grgl_dtype dt, grgl_direction dir, const std::vector<TravRequest>& reqs,
                      uint32_t Ktotal, const TravOptions& opt, char* rows, hipStream_t stream) {
    checkChainHealth(g);
    const TypeClass tc = typeClassOf(dt);
    const size_t es = dtypeSize(dt);
    const RowGeometry geo = rowGeometry(Ktotal, es);
    DeviceDirPlan& plan = dir == GRGL_DIRECTION_UP ? g->up : g->down;
    const bool hasInit = opt.initMode != GRGL_INIT_ZERO;
    // 1. Node initialisation (grg.h:1369-1402).
    if (hasInit) {
        for (const auto& r : reqs) {
            if (opt.initMode == GRGL_INIT_MATRIX) {
                // init is K x numNodes; row j takes column order[j].
                seedRows(tc, es, opt.init, g->V, r.K, rows, geo.ld, 0, r.kOffset, g->V, 1, plan.order, false, stream);
            } else {
                InitArgs a;
                a.rows = rows;
                a.ld = geo.ld;
                a.rowBegin = 0;
                a.rowEnd = g->V;
                a.kOffset = r.kOffset;
                a.K = r.K;
                a.kind = opt.initMode == GRGL_INIT_VECTOR ? kInitVector : kInitXtx;
                a.vec = opt.init;
                a.coalByRow = plan.coalByRow;
                launchInit(tc, a, stream);
            }
        }
    } else if (dir == GRGL_DIRECTION_UP && plan.levelStart[1] > g->N) {
        // Childless non-sample nodes of level 0 hold zero.
        InitArgs a{};
        a.rows = rows;
        a.ld = geo.ld;
        a.rowBegin = g->N;
        a.rowEnd = plan.levelStart[1];
        a.kOffset = 0;
        a.K = Ktotal;
        a.kind = kInitZero;
        launchInit(tc, a, stream);
    }
    // 2. Seeds.  Several requests without node initialisation: one stacked launch (and one
    // zero fill of the missingness seed rows for all columns).
    const std::vector<TravRequest> stacked = stackable(reqs);
    const bool stackSeeds = !stacked.empty() && !hasInit && (dir == GRGL_DIRECTION_UP || !opt.miss);
    auto launchSeeds = [&]() {
        if (stackSeeds) {
            if (dir == GRGL_DIRECTION_UP) {
                const uint32_t per = opt.byIndividual ? g->P : 1;
                const uint64_t inCols = opt.byIndividual ? g->N / g->P : g->N;
                stackedTranspose(tc, es, true, stacked, false, geo, rows, 0, g->N, inCols, per, nullptr, false, stream);
            } else {
                stackedTranspose(tc, es, true, stacked, false, geo, rows, g->V, g->M, g->M, 1, nullptr, false, stream);
                if (g->hasMiss) {
                    InitArgs a{};
                    a.rows = rows;
                    a.ld = geo.ld;
                    a.rowBegin = g->V + g->M;
                    a.rowEnd = g->V + 2 * g->M;
                    a.kOffset = 0;
                    a.K = (uint32_t)(geo.rowBytes / es);
                    a.kind = kInitZero;
                    launchInit(tc, a, stream);
                }
            }
        }
        for (const auto& r : reqs) {
            if (stackSeeds) break;
            if (dir == GRGL_DIRECTION_UP) {
                // Samples are rows 0..N-1 (grg.h:1456-1469).
                const uint32_t per = opt.byIndividual ? g->P : 1;
                const uint64_t inCols = opt.byIndividual ? g->N / g->P : g->N;
                seedRows(tc, es, r.in, inCols, r.K, rows, geo.ld, 0, r.kOffset, g->N, per, nullptr, hasInit, stream);
            } else {
                // Mutation seed rows V..V+M-1; missingness seed rows V+M..V+2M-1.
                seedRows(tc, es, r.in, g->M, r.K, rows, geo.ld, g->V, r.kOffset, g->M, 1, nullptr, false, stream);
                if (g->hasMiss) {
                    if (opt.miss) {
                        seedRows(tc, es, opt.miss, g->M, r.K, rows, geo.ld, g->V + g->M, r.kOffset, g->M, 1, nullptr,
                                 false, stream);
                    } else {
                        InitArgs a{};
                        a.rows = rows;
                        a.ld = geo.ld;
                        a.rowBegin = g->V + g->M;
                        a.rowEnd = g->V + 2 * g->M;
                        a.kOffset = r.kOffset;
                        a.K = r.K;
                        a.kind = kInitZero;
                        launchInit(tc, a, stream);
                    }
                }
            }
        }
        GRGL_HIP_CHECK(hipGetLastError());
    };
    // A one-column request of a 4- or 8-byte type without node initialisation, on a plan whose
    // first pulled level reads only seed rows (UP: level 1 reads only the samples, which are all
    // of level 0; DOWN: level 0, the roots, reads only mutation seeds, with no missingness rows):
    // the input row is the seed rows' layout, so the first level reads it there and its launch
    // copies it for the later levels (runLevels; the seed launch runs only where that does not
    // apply).  Saves the seed launch of a K = 1 call (4.5-4.8 us on the 200k x 1M graph).
    SeedFold sf{};
    const bool foldable = reqs.size() == 1 && Ktotal == 1 && reqs[0].kOffset == 0 && !hasInit && es >= 4 &&
                          geo.ld == es && plan.firstLevelSeedsOnly &&
                          (dir == GRGL_DIRECTION_UP ? !opt.byIndividual : !g->hasMiss);
    if (foldable) {
        sf.in = static_cast<const char*>(reqs[0].in);
        sf.seedBegin = dir == GRGL_DIRECTION_UP ? 0 : g->V;
        sf.bytes = (uint64_t)(dir == GRGL_DIRECTION_UP ? g->N : g->M) * es;
    } else {
        launchSeeds();
    }
    // 3. Levels.
    const Tuning tuning = chooseTuning(g, dir, tc, Ktotal, geo);
    // ORDER_PRESERVING: every list is summed by one team front to back, i.e. in the
    // reference's order.
    // DOWN traversals of one request with rows of at most 16 bytes (FP32 K <= 4) whose output
    // is the samples' rows as they are (no accumulation, no per-individual sums): the last
    // level, which holds every sample and which no row reads, writes the samples' values
    // straight to their output columns (0 + value, as the read-out does), and there is no
    // read-out launch.  On the 200k x 1M graph that is the sample level and saves the 4-5 us
    // read-out of a FP32 K = 1 call (0.112 -> 0.107 ms).
    OutFusion fo{};
    const bool fuse = dir == GRGL_DIRECTION_DOWN && reqs.size() == 1 && reqs[0].kOffset == 0 && geo.rowBytes <= 16 &&
                      !opt.emitAll && !opt.byIndividual && !opt.accumulate && plan.samplesLast;
    if (fuse) {
        fo.out = static_cast<char*>(reqs[0].out);
        fo.outCol = plan.order;
        fo.outCols = g->N;
        fo.K = reqs[0].K;
    }
    const bool fusedOut =
        runLevels(g, dir, tc, geo, tuning, hasInit, opt.orderPreserving, Ktotal, es, rows, stream, opt.launchOnly,
                  opt.chainState, fuse ? &fo : nullptr, foldable ? &sf : nullptr,
                  foldable ? std::function<void()>(launchSeeds) : std::function<void()>());
    if (fusedOut) {
        GRGL_HIP_CHECK(hipGetLastError());
        return;
    }
    // 4. Read-out.  Several requests: one stacked launch for the main outputs.
    const bool stackOut = !stacked.empty() && !(dir == GRGL_DIRECTION_UP && opt.miss && plan.missRow && !opt.emitAll);
    if (stackOut) {
        if (opt.emitAll) {
            stackedTranspose(tc, es, false, stacked, true, geo, rows, 0, g->V, g->V, 1, plan.rank, opt.accumulate,
                             stream);
        } else if (dir == GRGL_DIRECTION_UP) {
            stackedTranspose(tc, es, false, stacked, true, geo, rows, 0, g->M, g->M, 1, plan.outRow, opt.accumulate,
                             stream);
        } else {
            const uint32_t per = opt.byIndividual ? g->P : 1;
            const uint64_t C = g->N / per;
            stackedTranspose(tc, es, false, stacked, true, geo, rows, 0, C, C, per, plan.outRow, opt.accumulate,
                             stream);
        }
    }
    for (const auto& r : reqs) {
        if (stackOut) break;
        if (opt.emitAll) {
            readoutRows(tc, es, rows, geo.ld, r.kOffset, r.K, plan.rank, 1, g->V, r.out, g->V, opt.accumulate, stream);
        } else if (dir == GRGL_DIRECTION_UP) {
            readoutRows(tc, es, rows, geo.ld, r.kOffset, r.K, plan.outRow, 1, g->M, r.out, g->M, opt.accumulate, stream);
            if (opt.miss && plan.missRow) {
                readoutRows(tc, es, rows, geo.ld, r.kOffset, r.K, plan.missRow, 1, g->M, opt.miss, g->M, opt.accumulate,
                            stream);
            }
        } else {
            const uint32_t per = opt.byIndividual ? g->P : 1;
            const uint64_t C = g->N / per;
            readoutRows(tc, es, rows, geo.ld, r.kOffset, r.K, plan.outRow, per, C, r.out, C, opt.accumulate, stream);
        }
    }
    GRGL_HIP_CHECK(hipGetLastError());
}

Workspace* acquireWorkspace(grgl_graph* g, size_t bytes, hipStream_t stream) {
    std::unique_ptr<Workspace> victim; // freed outside the lock: hipFree synchronises the device
    Workspace* pick = nullptr;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        for (auto& w : g->workspaces) {
            if (w->busy || w->bytes < bytes) continue;
            if (!w->used || w->stream == stream || hipEventQuery(w->done) == hipSuccess) {
                pick = w.get();
                break;
            }
        }
        if (pick) {
            pick->busy = true;
            return pick;
        }
        // Grow: take out an idle, finished workspace that is too small, then allocate.
        for (auto it = g->workspaces.begin(); it != g->workspaces.end(); ++it) {
            Workspace& w = **it;
            if (!w.busy && w.bytes < bytes && (!w.used || hipEventQuery(w.done) == hipSuccess)) {
                victim = std::move(*it);
                g->workspaces.erase(it);
                break;
            }
        }
    }
    if (victim) {
        (void)hipFree(victim->raw);
        if (victim->done) (void)hipEventDestroy(victim->done);
    }
    std::unique_ptr<Workspace> w(new Workspace());
    const size_t sz = std::max<size_t>(bytes, 1 << 20);
    GRGL_HIP_CHECK(hipMalloc(&w->raw, sz + kChainStateBytes));
    w->chainState = reinterpret_cast<uint32_t*>(w->raw);
    w->ptr = w->raw + kChainStateBytes;
    w->bytes = sz;
    hipError_t ev = hipMemset(w->raw, 0, kChainStateBytes);
    if (ev == hipSuccess) ev = hipEventCreateWithFlags(&w->done, hipEventDisableTiming);
    if (ev != hipSuccess) {
        (void)hipFree(w->raw);
        GRGL_HIP_CHECK(ev);
    }
    w->busy = true;
    pick = w.get();
    std::lock_guard<std::mutex> lk(g->mu);
    g->workspaces.push_back(std::move(w));
    return pick;
}

void releaseWorkspace(grgl_graph* g, Workspace* ws, hipStream_t stream) {
    std::lock_guard<std::mutex> lk(g->mu);
    (void)hipEventRecord(ws->done, stream);
    ws->stream = stream;
    ws->used = true;
    ws->busy = false;
}

hipStream_t acquireInternalStream(grgl_graph* g) {
    std::lock_guard<std::mutex> lk(g->mu);
    if (!g->freeStreams.empty()) {
        hipStream_t s = g->freeStreams.back();
        g->freeStreams.pop_back();
        return s;
    }
    hipStream_t s;
    GRGL_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

void releaseInternalStream(grgl_graph* g, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g->mu);
    g->freeStreams.push_back(s);
}

} // namespace grgl_amd

using namespace grgl_amd;

namespace {

grgl_status translate(const std::exception& e) { return grgl_amd::statusOf(e); }

} // namespace

namespace grgl_amd {

void validateMatmulArgs(const grgl_graph* g, grgl_dtype dtype, const void* input, size_t input_cols, size_t input_rows,
                        grgl_direction direction, const void* output, size_t output_size, int emit_all_nodes,
                        int by_individual, const void* init, grgl_init_mode init_mode, const void* miss) {
    if (!g) throw MisuseError("null graph");
    const TypeClass tc = typeClassOf(dtype);
    validateShapes(g, direction, input_cols, input_rows, output_size, emit_all_nodes != 0, by_individual != 0);
    if (!input || !output) throw MisuseError("null input or output");
    if (tc == kX8 && init_mode != GRGL_INIT_ZERO)
        throw MisuseError("Node initialization and bit vector type cannot be mixed");
    if (init_mode == GRGL_INIT_XTX && !g->hasCoal)
        throw MisuseError("Coalescent counts not available for this GRG (init=XTX won't work)");
    if ((init_mode == GRGL_INIT_VECTOR || init_mode == GRGL_INIT_MATRIX) && !init)
        throw MisuseError("init mode requires an init buffer");
    if (miss && emit_all_nodes) throw MisuseError("The \"miss\" parameter cannot be mixed with the \"emit_all_nodes\" parameter");
}

grgl_status statusOf(const std::exception& e) {
    setLastError(e.what());
    if (dynamic_cast<const MisuseError*>(&e)) return GRGL_ERR_MISUSE;
    if (dynamic_cast<const NoMemError*>(&e)) return GRGL_ERR_NOMEM;
    if (dynamic_cast<const UnsupportedError*>(&e)) return GRGL_ERR_UNSUPPORTED;
    if (dynamic_cast<const DeviceError*>(&e)) return GRGL_ERR_DEVICE;
    if (dynamic_cast<const std::bad_alloc*>(&e)) return GRGL_ERR_NOMEM;
    return GRGL_ERR_MISUSE;
}

} // namespace grgl_amd

namespace {

#define GRGL_API_BEGIN try {
#define GRGL_API_END                                                                                   \
    return GRGL_OK;                                                                                    \
    }                                                                                                  \
    catch (const std::exception& e) {                                                                  \
        return translate(e);                                                                           \
    }

void uploadPlan(DeviceDirPlan& d, const DirPlan& p, uint64_t& bytes) {
    d.numNodes = p.numNodes;
    d.numLevels = p.numLevels;
    d.firstPulledLevel = p.firstPulledLevel;
    d.maxFanIn = p.maxFanIn;
    d.numEntries = p.src.size();
    d.levelStart = p.levelStart;
    d.fanIn.resize(p.numNodes);
    for (uint32_t j = 0; j < p.numNodes; j++) d.fanIn[j] = p.off[j + 1] - p.off[j];
    d.off = uploadVec(p.off, bytes);
    d.src = uploadVec(p.src, bytes);
    d.order = uploadVec(p.order, bytes);
    d.rank = uploadVec(p.rank, bytes);
    d.outRow = uploadVec(p.outRow, bytes);
    d.missRow = uploadVec(p.missRow, bytes);
    d.coalByRow = uploadVec(p.coalByRow, bytes);
    d.table = d.buildTable(kDefaultHeavyFanin, kChunkEntries);
}

// Grouped-launch plans of both directions (built on two host threads).  DOWN gets one more
// after its five classes: 64-byte slots over its levels of at most kLateLevelRows rows.
void buildGroups(const LevelizedGraph& lg, std::vector<GroupPlan>& up, std::vector<GroupPlan>& down) {
    const GroupLimits lim;
    std::exception_ptr err;
    std::thread t([&] {
        try {
            down = buildGroupPlans(lg.down, groupRegionEnd(lg.down, true), kDefaultHeavyFanin, lim);
            GroupLimits late;
            late.maxLevelRows = kLateLevelRows;
            late.classMask = 1U << (kGroupClasses - 1);
            down.push_back(
                buildGroupPlans(lg.down, groupRegionEnd(lg.down, true), kDefaultHeavyFanin, late)[kGroupClasses - 1]);
        } catch (...) {
            err = std::current_exception();
        }
    });
    up = buildGroupPlans(lg.up, groupRegionEnd(lg.up, false), kDefaultHeavyFanin, lim);
    t.join();
    if (err) std::rethrow_exception(err);
}

// The arrays of a grouped-launch plan, in the order of its device allocation.
std::vector<const std::vector<uint32_t>*> groupParts(const GroupPlan& gp) {
    return {&gp.blockRow, &gp.blockSeg, &gp.blockHeavy, &gp.blockImport, &gp.segRow,
            &gp.segHeavy, &gp.segHeavyOrd, &gp.row, &gp.off, &gp.pre,
            &gp.src, &gp.remOff, &gp.rem, &gp.heavyRow, &gp.importRow};
}

// Device bytes of a grouped-launch plan (each array padded to 16 bytes).
uint64_t groupDeviceBytes(const GroupPlan& gp) {
    if (gp.empty()) return 0;
    uint64_t words = 0;
    for (const auto* v : groupParts(gp)) words += (v->size() + 3) / 4 * 4;
    return 4 * words;
}

void uploadGroups(DeviceDirPlan& d, const std::vector<GroupPlan>& plans, uint64_t& bytes) {
    for (uint32_t c = 0; c <= kDeviceGroupClasses && c < plans.size(); c++) {
        const GroupPlan& gp = plans[c];
        DeviceGroupPlan& dg = c < kDeviceGroupClasses ? d.groups[c] : d.lateWide;
        dg = DeviceGroupPlan();
        dg.slotBytes = gp.slotBytes;
        dg.heavyFanin = gp.heavyFanin;
        if (gp.empty()) continue;
        const std::vector<const std::vector<uint32_t>*> parts = groupParts(gp);
        std::vector<uint32_t> flat;
        std::vector<size_t> at;
        for (const auto* v : parts) {
            at.push_back(flat.size());
            flat.insert(flat.end(), v->begin(), v->end());
            flat.resize((flat.size() + 3) / 4 * 4); // 16-byte aligned parts
        }
        dg.mem = uploadVec(flat, bytes);
        const uint32_t** dst[] = {&dg.blockRow, &dg.blockSeg, &dg.blockHeavy, &dg.blockImport, &dg.segRow,
                                  &dg.segHeavy, &dg.segHeavyOrd, &dg.row, &dg.off, &dg.pre, &dg.src, &dg.remOff,
                                  &dg.rem, &dg.heavyRow, &dg.importRow};
        for (size_t i = 0; i < parts.size(); i++) *dst[i] = dg.mem + at[i];
        dg.levelBegin = gp.levelBegin;
        dg.levelEnd = gp.levelEnd;
        dg.blocks = gp.blocks;
        dg.ldsBytes = gp.ldsBytes;
        dg.numRows = (uint32_t)gp.row.size();
        dg.numHeavy = (uint32_t)gp.heavyRow.size();
        dg.prefixEntries = 0;
        for (uint32_t x : gp.pre) dg.prefixEntries += x;
    }
}

void freePlan(DeviceDirPlan& d) {
    freeDev(d.off);
    freeDev(d.src);
    freeDev(d.order);
    freeDev(d.rank);
    freeDev(d.outRow);
    freeDev(d.missRow);
    freeDev(d.coalByRow);
    for (auto& dg : d.groups) freeDev(dg.mem);
    freeDev(d.lateWide.mem);
    d = DeviceDirPlan();
}

// The host form of a graph descriptor, validated (misuse errors for bad input).
HostGrg hostGrgOf(const grgl_graph_desc& desc) {
    HostGrg h;
    h.numSamples = desc.num_samples;
    h.ploidy = desc.ploidy;
    if (!desc.down_offsets) throw MisuseError("down_offsets is required");
    h.downOff.assign(desc.down_offsets, desc.down_offsets + (size_t)desc.num_nodes + 1);
    if (h.downOff.back() > 0 && !desc.down_children) throw MisuseError("down_children is required");
    if (desc.num_mutations > 0 && !desc.mutation_nodes) throw MisuseError("mutation_nodes is required");
    if (h.downOff.back() > 0) h.downChild.assign(desc.down_children, desc.down_children + h.downOff.back());
    if (desc.num_mutations > 0) h.mutNode.assign(desc.mutation_nodes, desc.mutation_nodes + desc.num_mutations);
    if (desc.mutation_miss_nodes) {
        h.mutMiss.assign(desc.mutation_miss_nodes, desc.mutation_miss_nodes + desc.num_mutations);
        bool any = false;
        for (uint32_t x : h.mutMiss) any |= (x != kInvalidNode);
        if (!any) h.mutMiss.clear();
    }
    if (desc.coal_counts) h.coal.assign(desc.coal_counts, desc.coal_counts + desc.num_nodes);
    try {
        validateHostGrg(h);
    } catch (const std::runtime_error& e) {
        throw MisuseError(e.what());
    }
    return h;
}

// Whether every gather entry of level `lvl` of `p` reads a row in [lo, hi).
bool levelReadsOnly(const DirPlan& p, uint32_t lvl, uint64_t lo, uint64_t hi) {
    uint64_t pos = 0;
    for (uint32_t r = 0; r < p.levelStart[lvl]; r++) pos += (uint32_t)(p.off[r + 1] - p.off[r]);
    for (uint32_t r = p.levelStart[lvl]; r < p.levelStart[lvl + 1]; r++) {
        const uint64_t end = pos + (uint32_t)(p.off[r + 1] - p.off[r]);
        for (; pos < end; pos++)
            if (p.src[pos] < lo || p.src[pos] >= hi) return false;
    }
    return true;
}

// A device graph from both plans: upload, control words, clock rate.
std::unique_ptr<grgl_graph> makeDeviceGraph(LevelizedGraph lg, int device) {
    int ndev = 0;
    GRGL_HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw MisuseError("invalid device ordinal " + std::to_string(device));
    GRGL_HIP_CHECK(hipSetDevice(device));
    std::unique_ptr<grgl_graph> g(new grgl_graph());
    g->device = device;
    g->N = lg.N;
    g->V = lg.V;
    g->M = lg.M;
    g->P = lg.P;
    g->E = lg.E;
    g->hasMiss = lg.hasMiss;
    g->hasCoal = lg.hasCoal;
    try {
        std::vector<GroupPlan> upGroups, downGroups;
        buildGroups(lg, upGroups, downGroups);
        uploadPlan(g->up, lg.up, g->deviceBytes);
        uploadPlan(g->down, lg.down, g->deviceBytes);
        g->down.down = true;
        // Every sample in the last DOWN level (samples have no children, so always in a GRG):
        // the condition for writing that level's rows straight to the output.
        g->down.samplesLast = lg.down.numLevels > 0;
        for (uint32_t sIdx = 0; sIdx < lg.N && g->down.samplesLast; sIdx++)
            g->down.samplesLast = lg.down.outRow[sIdx] >= lg.down.levelStart[lg.down.numLevels - 1];
        // First pulled level reading only seed rows (UP: the samples, all of level 0; DOWN:
        // mutation seeds): the condition for reading a one-column input in place.
        g->up.firstLevelSeedsOnly = lg.up.numLevels > 1 && lg.up.levelStart[1] == lg.N &&
                                    levelReadsOnly(lg.up, 1, 0, lg.N);
        g->down.firstLevelSeedsOnly = lg.down.numLevels > 0 && levelReadsOnly(lg.down, 0, lg.V, lg.V + lg.M);
        uploadGroups(g->up, upGroups, g->deviceBytes);
        uploadGroups(g->down, downGroups, g->deviceBytes);
        void* ctl = nullptr;
        GRGL_HIP_CHECK(hipHostMalloc(&ctl, sizeof(ChainControl), hipHostMallocMapped));
        g->chainCtl = static_cast<ChainControl*>(ctl);
        std::memset(g->chainCtl, 0, sizeof(ChainControl));
        void* dctl = nullptr;
        GRGL_HIP_CHECK(hipHostGetDevicePointer(&dctl, ctl, 0));
        g->chainCtlDev = static_cast<ChainControl*>(dctl);
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            g->wallClockKHz = (uint64_t)khz;
    } catch (...) {
        freePlan(g->up);
        freePlan(g->down);
        if (g->chainCtl) (void)hipHostFree(g->chainCtl);
        throw;
    }
    return g;
}

template <typename T> void downloadVec(const T* d, size_t n, std::vector<T>& v) {
    v.resize(n);
    if (n) GRGL_HIP_CHECK(hipMemcpy(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
}

// The host plan of a device plan (the arrays uploadPlan copied; outRows and missRows as
// counted by the direction).
void downloadPlan(const DeviceDirPlan& d, DirPlan& p, uint32_t V, uint32_t outRows, bool miss) {
    p.numNodes = d.numNodes;
    p.numLevels = d.numLevels;
    p.firstPulledLevel = d.firstPulledLevel;
    p.maxFanIn = d.maxFanIn;
    p.levelStart = d.levelStart;
    downloadVec(d.off, (size_t)V + 1, p.off);
    downloadVec(d.src, d.numEntries, p.src);
    downloadVec(d.order, V, p.order);
    downloadVec(d.rank, V, p.rank);
    downloadVec(d.outRow, outRows, p.outRow);
    downloadVec(d.missRow, miss ? outRows : 0, p.missRow);
    downloadVec(d.coalByRow, V, p.coalByRow);
}

} // namespace

GRGL_EXPORT const char* grgl_hip_last_error(void) { return g_lastError.c_str(); }

GRGL_EXPORT grgl_status grgl_hip_host_alloc(size_t bytes, void** out) {
    GRGL_API_BEGIN
    if (!out) throw MisuseError("null argument");
    *out = nullptr;
    if (bytes == 0) throw MisuseError("zero-byte allocation");
    GRGL_HIP_CHECK(hipHostMalloc(out, bytes, hipHostMallocDefault));
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_host_free(void* ptr) {
    GRGL_API_BEGIN
    if (ptr) GRGL_HIP_CHECK(hipHostFree(ptr));
    GRGL_API_END
}

GRGL_EXPORT int grgl_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

GRGL_EXPORT grgl_status grgl_hip_graph_create(const grgl_graph_desc* desc, int device, grgl_graph** out) {
    GRGL_API_BEGIN
    if (!desc || !out) throw MisuseError("null argument");
    *out = nullptr;
    HostGrg h = hostGrgOf(*desc);
    if ((uint64_t)h.numNodes() + 2ULL * h.numMutations() >= 0x7fffffffULL)
        throw MisuseError("graph too large: nodes + 2*mutations must stay below 2^31");
    *out = makeDeviceGraph(levelize(h), device).release();
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_levelize_to_file(const grgl_graph_desc* desc, const char* path) {
    GRGL_API_BEGIN
    if (!desc || !path) throw MisuseError("null argument");
    HostGrg h = hostGrgOf(*desc);
    if ((uint64_t)h.numNodes() + 2ULL * h.numMutations() >= 0x7fffffffULL)
        throw MisuseError("graph too large: nodes + 2*mutations must stay below 2^31");
    writeLevelizedFile(levelize(h), path);
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_graph_store(const grgl_graph* g, const char* path) {
    GRGL_API_BEGIN
    if (!g || !path) throw MisuseError("null argument");
    GRGL_HIP_CHECK(hipSetDevice(g->device));
    LevelizedGraph lg;
    lg.N = g->N;
    lg.V = g->V;
    lg.M = g->M;
    lg.P = g->P;
    lg.E = g->E;
    lg.hasMiss = g->hasMiss;
    lg.hasCoal = g->hasCoal;
    downloadPlan(g->up, lg.up, lg.V, lg.M, lg.hasMiss);
    downloadPlan(g->down, lg.down, lg.V, lg.N, false);
    writeLevelizedFile(lg, path);
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_graph_load(const char* path, int device, grgl_graph** out) {
    GRGL_API_BEGIN
    if (!path || !out) throw MisuseError("null argument");
    *out = nullptr;
    LevelizedGraph lg;
    try {
        lg = readLevelizedFile(path);
    } catch (const std::runtime_error& e) {
        throw MisuseError(e.what());
    }
    *out = makeDeviceGraph(std::move(lg), device).release();
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_levelized_file_info(const char* path, grgl_graph_info* info) {
    GRGL_API_BEGIN
    if (!path || !info) throw MisuseError("null argument");
    LevelizedGraph lg;
    try {
        lg = readLevelizedFile(path);
    } catch (const std::runtime_error& e) {
        throw MisuseError(e.what());
    }
    *info = grgl_graph_info{};
    info->num_samples = lg.N;
    info->num_nodes = lg.V;
    info->num_mutations = lg.M;
    info->ploidy = lg.P;
    info->num_edges = lg.E;
    info->up_levels = lg.up.numLevels;
    info->down_levels = lg.down.numLevels;
    info->max_fanin_up = lg.up.maxFanIn;
    info->max_fanin_down = lg.down.maxFanIn;
    info->up_gather_entries = lg.up.src.size();
    info->down_gather_entries = lg.down.src.size();
    uint64_t bytes = 0;
    for (const DirPlan* p : {&lg.up, &lg.down})
        bytes += 4 * (p->off.size() + p->src.size() + p->order.size() + p->rank.size() + p->outRow.size() +
                      p->missRow.size() + p->coalByRow.size());
    std::vector<GroupPlan> upGroups, downGroups;
    buildGroups(lg, upGroups, downGroups);
    for (const auto* plans : {&upGroups, &downGroups})
        for (const GroupPlan& gp : *plans) bytes += groupDeviceBytes(gp);
    info->device_bytes = bytes;
    info->device = -1;
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_graph_destroy(grgl_graph* g) {
    GRGL_API_BEGIN
    if (!g) return GRGL_OK;
    (void)hipSetDevice(g->device);
    (void)hipDeviceSynchronize();
    g->graphs.clear(); // CapturedLevels destroys its graph, exec and event
    for (auto& w : g->workspaces) {
        (void)hipFree(w->raw);
        if (w->done) (void)hipEventDestroy(w->done);
    }
    for (auto s : g->freeStreams) (void)hipStreamDestroy(s);
    for (auto& ev : g->timingEvents) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    freePlan(g->up);
    freePlan(g->down);
    if (g->chainCtl) (void)hipHostFree(g->chainCtl);
    delete g;
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_graph_info(const grgl_graph* g, grgl_graph_info* info) {
    GRGL_API_BEGIN
    if (!g || !info) throw MisuseError("null argument");
    info->num_samples = g->N;
    info->num_nodes = g->V;
    info->num_mutations = g->M;
    info->ploidy = g->P;
    info->num_edges = g->E;
    info->up_levels = g->up.numLevels;
    info->down_levels = g->down.numLevels;
    info->max_fanin_up = g->up.maxFanIn;
    info->max_fanin_down = g->down.maxFanIn;
    info->up_gather_entries = g->up.numEntries;
    info->down_gather_entries = g->down.numEntries;
    info->device_bytes = g->deviceBytes;
    info->device = g->device;
    GRGL_API_END
}

namespace {

// Copy between host and device memory where the host side may be pageable.  The runtime
// stages pageable copies through page-locked buffers in the calling thread, so a large
// pageable buffer is split over 2 or 4 threads, each copying its slice on an internal stream
// that first waits for the work already on `s`; `s` then waits for every slice.
bool pageLocked(const void* host) {
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, host) != hipSuccess) {
        (void)hipGetLastError(); // pageable memory is unknown to the runtime
        return false;
    }
    return attr.type == hipMemoryTypeHost;
}

void copyHostSplit(grgl_graph* g, void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
    constexpr size_t kSlice = 16ULL << 20;
    unsigned T = bytes >= 4 * kSlice ? 4 : (bytes >= 2 * kSlice ? 2 : 1);
    // Page-locked host memory is copied directly at link speed: no split.
    if (T > 1 && pageLocked(kind == hipMemcpyHostToDevice ? src : dst)) T = 1;
    if (T == 1) {
        GRGL_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, kind, s));
        return;
    }
    hipEvent_t before = nullptr;
    GRGL_HIP_CHECK(hipEventCreateWithFlags(&before, hipEventDisableTiming));
    std::vector<hipStream_t> st(T, nullptr);
    std::vector<hipEvent_t> ev(T, nullptr);
    std::vector<hipError_t> err(T, hipSuccess);
    hipError_t first = hipEventRecord(before, s);
    for (unsigned t = 0; t < T && first == hipSuccess; t++) {
        st[t] = acquireInternalStream(g);
        first = hipEventCreateWithFlags(&ev[t], hipEventDisableTiming);
        if (first == hipSuccess) first = hipStreamWaitEvent(st[t], before, 0);
    }
    if (first == hipSuccess) {
        std::vector<std::thread> threads;
        for (unsigned t = 0; t < T; t++) {
            threads.emplace_back([&, t] {
                const size_t lo = bytes * t / T, hi = bytes * (t + 1) / T;
                hipError_t e = hipSetDevice(g->device);
                if (e == hipSuccess) e = hipMemcpyAsync((char*)dst + lo, (const char*)src + lo, hi - lo, kind, st[t]);
                if (e == hipSuccess) e = hipEventRecord(ev[t], st[t]);
                err[t] = e;
            });
        }
        for (auto& th : threads) th.join();
        for (unsigned t = 0; t < T; t++) {
            if (err[t] != hipSuccess && first == hipSuccess) first = err[t];
            if (err[t] == hipSuccess) {
                const hipError_t w = hipStreamWaitEvent(s, ev[t], 0);
                if (w != hipSuccess && first == hipSuccess) first = w;
            }
        }
    }
    for (unsigned t = 0; t < T; t++) {
        if (ev[t]) (void)hipEventDestroy(ev[t]);
        if (st[t]) releaseInternalStream(g, st[t]);
    }
    (void)hipEventDestroy(before);
    GRGL_HIP_CHECK(first);
}

// One traversal of up to 65536 input rows (grgl_hip_matmul splits larger inputs).  Shapes and
// arguments are validated by the caller.
void matmulRows(grgl_graph* g, grgl_dtype dtype, const void* input, size_t input_cols, uint32_t K,
                grgl_direction direction, void* output, size_t output_size, int emit_all_nodes, int by_individual,
                const void* init, grgl_init_mode init_mode, void* miss, uint32_t flags, void* stream) {
    const size_t input_rows = K;
    const size_t es = dtypeSize(dtype);
    const bool host = (flags & GRGL_FLAG_HOST_PTRS) != 0;
    hipStream_t s = (hipStream_t)stream;
    bool ownStream = false;
    if (host && s == nullptr) {
        s = acquireInternalStream(g);
        ownStream = true;
    }
    const size_t rowsBytes = rowsBufferBytes(g, direction, K, es);
    auto align = [](size_t x) { return (x + 255) / 256 * 256; };
    size_t total = align(rowsBytes);
    size_t inOff = 0, outOff = 0, initOff = 0, missOff = 0;
    const size_t inBytes = input_rows * input_cols * es, outBytes = output_size * es;
    const size_t initBytes = init_mode == GRGL_INIT_VECTOR ? K * es : (init_mode == GRGL_INIT_MATRIX ? (size_t)K * g->V * es : 0);
    const size_t missBytes = miss ? (size_t)K * g->M * es : 0;
    if (host) {
        inOff = total;
        total += align(inBytes);
        outOff = total;
        total += align(outBytes);
        initOff = total;
        total += align(initBytes);
        missOff = total;
        total += align(missBytes);
    }
    Workspace* ws = acquireWorkspace(g, total, s);
    try {
        char* base = ws->ptr;
        const void* dIn = input;
        void* dOut = output;
        const void* dInit = init;
        void* dMiss = miss;
        const bool accumulate = (flags & GRGL_FLAG_ACCUMULATE) != 0;
        if (host) {
            copyHostSplit(g, base + inOff, input, inBytes, hipMemcpyHostToDevice, s);
            dIn = base + inOff;
            dOut = base + outOff;
            if (accumulate) copyHostSplit(g, dOut, output, outBytes, hipMemcpyHostToDevice, s);
            if (initBytes) {
                copyHostSplit(g, base + initOff, init, initBytes, hipMemcpyHostToDevice, s);
                dInit = base + initOff;
            }
            if (missBytes) {
                copyHostSplit(g, base + missOff, miss, missBytes, hipMemcpyHostToDevice, s);
                dMiss = base + missOff;
            }
        }
        TravOptions opt;
        opt.emitAll = emit_all_nodes != 0;
        opt.byIndividual = by_individual != 0;
        opt.initMode = init_mode;
        opt.init = dInit;
        opt.miss = dMiss;
        opt.accumulate = accumulate;
        opt.orderPreserving = (flags & GRGL_FLAG_ORDER_PRESERVING) != 0;
        opt.chainState = ws->chainState;
        std::vector<TravRequest> reqs{{K, dIn, dOut, 0}};
        enqueueTraversal(g, dtype, direction, reqs, K, opt, base, s);
        if (host) {
            copyHostSplit(g, output, dOut, outBytes, hipMemcpyDeviceToHost, s);
            if (missBytes && direction == GRGL_DIRECTION_UP)
                copyHostSplit(g, miss, dMiss, missBytes, hipMemcpyDeviceToHost, s);
            GRGL_HIP_CHECK(hipStreamSynchronize(s));
            checkChainHealth(g);
        }
    } catch (...) {
        releaseWorkspace(g, ws, s);
        if (ownStream) releaseInternalStream(g, s);
        throw;
    }
    releaseWorkspace(g, ws, s);
    if (ownStream) releaseInternalStream(g, s);
}

// Input rows per traversal: 65536, or fewer when the row buffer (plus staging of host
// buffers) would take more than half of the free device memory.  Rows are independent, so a
// larger input is computed block by block with identical results.
uint32_t rowsPerTraversal(grgl_graph* g, grgl_direction dir, size_t es, bool host, size_t inCols, size_t outCols,
                          grgl_init_mode initMode, bool hasMiss) {
    uint32_t kb = 65536;
    // Host buffers are staged in the same workspace as the row buffer: input, output, the
    // init vector or matrix and the miss matrix (as matmulRows lays them out).
    auto need = [&](uint32_t k) {
        size_t b = rowsBufferBytes(g, dir, k, es);
        if (host) {
            b += (size_t)k * (inCols + outCols) * es;
            if (initMode == GRGL_INIT_VECTOR) b += (size_t)k * es;
            else if (initMode == GRGL_INIT_MATRIX) b += (size_t)k * g->V * es;
            if (hasMiss) b += (size_t)k * g->M * es;
        }
        return b;
    };
    if (need(kb) <= (1ULL << 30)) return kb; // small graphs: no device query
    size_t freeBytes = 0, totalBytes = 0;
    if (hipMemGetInfo(&freeBytes, &totalBytes) != hipSuccess) return kb;
    while (kb > 1 && need(kb) > freeBytes / 2) kb /= 2;
    return kb;
}

} // namespace

GRGL_EXPORT grgl_status grgl_hip_matmul(grgl_graph* g, grgl_dtype dtype, const void* input, size_t input_cols,
                                        size_t input_rows, grgl_direction direction, void* output, size_t output_size,
                                        int emit_all_nodes, int by_individual, const void* init,
                                        grgl_init_mode init_mode, void* miss, uint32_t flags, void* stream) {
    GRGL_API_BEGIN
    validateMatmulArgs(g, dtype, input, input_cols, input_rows, direction, output, output_size, emit_all_nodes,
                       by_individual, init, init_mode, miss);
    GRGL_HIP_CHECK(hipSetDevice(g->device));
    const size_t es = dtypeSize(dtype);
    const size_t outCols = output_size / input_rows;
    const uint32_t kb = (uint32_t)std::min<size_t>(
        input_rows, rowsPerTraversal(g, direction, es, (flags & GRGL_FLAG_HOST_PTRS) != 0, input_cols, outCols,
                                     init_mode, miss != nullptr));
    // Row blocks: input, output, init and miss are row-major with one row per input row.
    for (size_t k0 = 0; k0 < input_rows; k0 += kb) {
        const uint32_t rows = (uint32_t)std::min<size_t>(kb, input_rows - k0);
        const void* blockInit = init;
        if (init_mode == GRGL_INIT_VECTOR) blockInit = (const char*)init + k0 * es;
        else if (init_mode == GRGL_INIT_MATRIX) blockInit = (const char*)init + k0 * g->V * es;
        matmulRows(g, dtype, (const char*)input + k0 * input_cols * es, input_cols, rows, direction,
                   (char*)output + k0 * outCols * es, (size_t)rows * outCols, emit_all_nodes, by_individual, blockInit,
                   init_mode, miss ? (char*)miss + k0 * g->M * es : nullptr, flags, stream);
    }
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_matmul_batch(grgl_graph* g, grgl_dtype dtype, grgl_direction direction,
                                              const grgl_request* requests, uint32_t num_requests, uint32_t flags,
                                              void* stream) {
    GRGL_API_BEGIN
    if (!g) throw MisuseError("null graph");
    if (num_requests == 0) return GRGL_OK;
    if (!requests) throw MisuseError("null requests");
    const size_t es = dtypeSize(dtype);
    const size_t inCols = direction == GRGL_DIRECTION_UP ? g->N : g->M;
    const size_t outCols = direction == GRGL_DIRECTION_UP ? g->M : g->N;
    std::vector<uint32_t> rowsOf(num_requests);
    for (uint32_t i = 0; i < num_requests; i++) {
        if (requests[i].rows == 0) throw MisuseError("matmul() requires non-zero dimensions.");
        if (!requests[i].input || !requests[i].output) throw MisuseError("null input or output");
        rowsOf[i] = requests[i].rows;
    }
    uint64_t Ktot = 0;
    const std::vector<uint32_t> offsets = stackOffsets(rowsOf, es, Ktot);
    if (direction != GRGL_DIRECTION_UP && direction != GRGL_DIRECTION_DOWN) throw MisuseError("Invalid traversal direction");
    if (Ktot > 65536) throw MisuseError("at most 65536 stacked rows per batch");
    const uint32_t K = (uint32_t)Ktot;
    GRGL_HIP_CHECK(hipSetDevice(g->device));
    const bool host = (flags & GRGL_FLAG_HOST_PTRS) != 0;
    hipStream_t s = (hipStream_t)stream;
    bool ownStream = false;
    if (host && s == nullptr) {
        s = acquireInternalStream(g);
        ownStream = true;
    }
    auto align = [](size_t x) { return (x + 255) / 256 * 256; };
    size_t total = align(rowsBufferBytes(g, direction, K, es));
    const size_t stageBase = total;
    if (host) total += align((size_t)K * inCols * es) + align((size_t)K * outCols * es);
    Workspace* ws = acquireWorkspace(g, total, s);
    try {
        char* base = ws->ptr;
        std::vector<TravRequest> reqs;
        uint32_t k0 = 0;
        char* inStage = base + stageBase;
        char* outStage = inStage + align((size_t)K * inCols * es);
        const bool accumulate = (flags & GRGL_FLAG_ACCUMULATE) != 0;
        for (uint32_t i = 0; i < num_requests; i++) {
            const auto& r = requests[i];
            k0 = offsets[i];
            TravRequest t{r.rows, r.input, r.output, k0};
            if (host) {
                char* in = inStage + (size_t)k0 * inCols * es;
                char* out = outStage + (size_t)k0 * outCols * es;
                GRGL_HIP_CHECK(hipMemcpyAsync(in, r.input, (size_t)r.rows * inCols * es, hipMemcpyHostToDevice, s));
                if (accumulate)
                    GRGL_HIP_CHECK(hipMemcpyAsync(out, r.output, (size_t)r.rows * outCols * es, hipMemcpyHostToDevice, s));
                t.in = in;
                t.out = out;
            }
            reqs.push_back(t);
        }
        TravOptions opt;
        opt.accumulate = accumulate;
        opt.orderPreserving = (flags & GRGL_FLAG_ORDER_PRESERVING) != 0;
        opt.chainState = ws->chainState;
        enqueueTraversal(g, dtype, direction, reqs, K, opt, base, s);
        if (host) {
            for (uint32_t i = 0; i < num_requests; i++) {
                GRGL_HIP_CHECK(hipMemcpyAsync(requests[i].output, reqs[i].out, (size_t)requests[i].rows * outCols * es,
                                              hipMemcpyDeviceToHost, s));
            }
            GRGL_HIP_CHECK(hipStreamSynchronize(s));
            checkChainHealth(g);
        }
    } catch (...) {
        releaseWorkspace(g, ws, s);
        if (ownStream) releaseInternalStream(g, s);
        throw;
    }
    releaseWorkspace(g, ws, s);
    if (ownStream) releaseInternalStream(g, s);
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_set_tuning(grgl_graph* g, grgl_direction direction, grgl_dtype dtype, uint32_t rows,
                                            const grgl_tuning* tuning) {
    GRGL_API_BEGIN
    if (!g || !tuning) throw MisuseError("null argument");
    if (direction != GRGL_DIRECTION_UP && direction != GRGL_DIRECTION_DOWN) throw MisuseError("Invalid traversal direction");
    const TypeClass tc = typeClassOf(dtype);
    if (tuning->team_lanes > 64) throw MisuseError("team_lanes must be at most 64");
    std::lock_guard<std::mutex> lk(g->mu);
    Tuning t;
    t.teamLanes = tuning->team_lanes;
    t.useGraph = tuning->use_graph;
    if (tuning->xcd_map > 2) throw MisuseError("xcd_map must be 0, 1 or 2");
    t.xcdMap = tuning->xcd_map;
    if (tuning->chain > 3) throw MisuseError("chain must be 0, 1, 2 or 3");
    t.chain = tuning->chain;
    if (tuning->block_threads != 0 && tuning->block_threads != 256 && tuning->block_threads != 512 &&
        tuning->block_threads != 1024)
        throw MisuseError("block_threads must be 0, 256, 512 or 1024");
    if (tuning->loads_in_flight != 0 && tuning->loads_in_flight != 4 && tuning->loads_in_flight != 8)
        throw MisuseError("loads_in_flight must be 0, 4 or 8");
    t.blockThreads = tuning->block_threads;
    t.loadsInFlight = tuning->loads_in_flight;
    g->tuning[std::make_tuple((int)direction, (int)tc, rows)] = t;
    if (tuning->heavy_fanin || tuning->max_launch_entries) {
        // A new launch table replaces the direction's current one; traversals already
        // running keep the snapshot they took.
        DeviceDirPlan& plan = direction == GRGL_DIRECTION_UP ? g->up : g->down;
        const uint32_t heavy = tuning->heavy_fanin ? tuning->heavy_fanin : plan.table->heavyFanin;
        const uint64_t chunk = tuning->max_launch_entries ? std::min<uint64_t>(tuning->max_launch_entries, kChunkEntries)
                                                          : plan.table->chunkEntries;
        if (heavy != plan.table->heavyFanin || chunk != plan.table->chunkEntries) plan.table = plan.buildTable(heavy, chunk);
    }
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_chain_control(grgl_graph* g, int abort, uint32_t wait_us, uint32_t* gave_up,
                                               int reset) {
    GRGL_API_BEGIN
    if (!g) throw MisuseError("null graph");
    if (abort < -1 || abort > 1) throw MisuseError("abort must be -1, 0 or 1");
    if (abort >= 0) __atomic_store_n(&g->chainCtl->abortWord, (uint32_t)abort, __ATOMIC_RELEASE);
    if (wait_us) g->chainWaitUs = wait_us;
    if (gave_up) *gave_up = __atomic_load_n(&g->chainCtl->gaveUp, __ATOMIC_ACQUIRE);
    if (reset) __atomic_store_n(&g->chainCtl->gaveUp, 0U, __ATOMIC_RELEASE);
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_get_tuning(grgl_graph* g, grgl_direction direction, grgl_dtype dtype, uint32_t rows,
                                            grgl_tuning* out) {
    GRGL_API_BEGIN
    if (!g || !out) throw MisuseError("null argument");
    if (direction != GRGL_DIRECTION_UP && direction != GRGL_DIRECTION_DOWN) throw MisuseError("Invalid traversal direction");
    if (rows == 0) throw MisuseError("rows must be positive");
    const TypeClass tc = typeClassOf(dtype);
    const Tuning t = chooseTuning(g, direction, tc, rows, rowGeometry(rows, dtypeSize(dtype)));
    std::lock_guard<std::mutex> lk(g->mu);
    const DeviceDirPlan& plan = direction == GRGL_DIRECTION_UP ? g->up : g->down;
    out->team_lanes = t.teamLanes;
    out->heavy_fanin = plan.table->heavyFanin;
    out->use_graph = t.useGraph;
    out->xcd_map = t.xcdMap;
    const RowGeometry geo = rowGeometry(rows, dtypeSize(dtype));
    tileShape(tc, levelVb(geo, t.teamLanes, dtypeSize(dtype)), t.teamLanes, t, out->block_threads, out->loads_in_flight);
    reportChain(plan, *plan.table, tc, geo, t, *out);
    out->max_launch_entries = (uint32_t)plan.table->chunkEntries;
    GRGL_API_END
}

namespace {

// Fingerprint of what tuning depends on: sizes, level boundaries and list lengths of both
// directions (FNV-1a, 64 bit).
uint64_t tuningFingerprint(const grgl_graph* g) {
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&h](uint64_t v) {
        for (int i = 0; i < 8; i++) {
            h ^= (v >> (8 * i)) & 0xff;
            h *= 1099511628211ULL;
        }
    };
    mix(g->V);
    mix(g->E);
    mix(g->N);
    mix(g->M);
    for (const DeviceDirPlan* p : {&g->up, &g->down}) {
        mix(p->numLevels);
        for (uint32_t x : p->levelStart) mix(x);
        for (uint32_t x : p->fanIn) mix(x);
    }
    return h;
}

const char* kTuningMagic = "grgl_hip_tuning";

} // namespace

GRGL_EXPORT grgl_status grgl_hip_tuning_save(grgl_graph* g, const char* path) {
    GRGL_API_BEGIN
    if (!g || !path) throw MisuseError("null argument");
    std::ostringstream os;
    os << kTuningMagic << " 3\n";
    char fp[32];
    std::snprintf(fp, sizeof(fp), "%016llx", (unsigned long long)tuningFingerprint(g));
    os << "graph " << fp << "\n";
    {
        std::lock_guard<std::mutex> lk(g->mu);
        os << "direction " << (int)GRGL_DIRECTION_UP << " " << g->up.table->heavyFanin << " "
           << g->up.table->chunkEntries << "\n";
        os << "direction " << (int)GRGL_DIRECTION_DOWN << " " << g->down.table->heavyFanin << " "
           << g->down.table->chunkEntries << "\n";
        for (const auto& kv : g->tuning) {
            const Tuning& t = kv.second;
            os << "entry " << std::get<0>(kv.first) << " " << std::get<1>(kv.first) << " " << std::get<2>(kv.first) << " "
               << t.teamLanes << " " << t.useGraph << " " << t.xcdMap << " " << t.chain << " " << t.blockThreads
               << " " << t.loadsInFlight << "\n";
        }
    }
    std::ofstream f(path, std::ios::trunc);
    if (!f) throw MisuseError(std::string("cannot write tuning file ") + path);
    f << os.str();
    if (!f.good()) throw MisuseError(std::string("cannot write tuning file ") + path);
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_tuning_load(grgl_graph* g, const char* path) {
    GRGL_API_BEGIN
    if (!g || !path) throw MisuseError("null argument");
    std::ifstream f(path);
    if (!f) throw MisuseError(std::string("cannot read tuning file ") + path);
    std::string magic, word, fp;
    int version = 0;
    // Version 1 files carry no chain column, versions 1 and 2 no block size and loads in
    // flight (their entries keep the defaults).
    if (!(f >> magic >> version) || magic != kTuningMagic || version < 1 || version > 3)
        throw MisuseError("not a tuning file");
    if (!(f >> word >> fp) || word != "graph") throw MisuseError("malformed tuning file");
    char want[32];
    std::snprintf(want, sizeof(want), "%016llx", (unsigned long long)tuningFingerprint(g));
    if (fp != want) throw MisuseError("tuning file was written for a graph of another structure");
    // Parse everything before applying anything.
    struct Dir { int dir; uint64_t heavy, chunk; };
    std::vector<Dir> dirs;
    std::vector<std::pair<std::tuple<int, int, uint32_t>, Tuning>> entries;
    while (f >> word) {
        if (word == "direction") {
            Dir d{};
            if (!(f >> d.dir >> d.heavy >> d.chunk) || (d.dir != GRGL_DIRECTION_UP && d.dir != GRGL_DIRECTION_DOWN) ||
                d.heavy == 0 || d.heavy > 0xffffffffULL || d.chunk == 0 || d.chunk > kChunkEntries)
                throw MisuseError("malformed tuning file (direction)");
            dirs.push_back(d);
        } else if (word == "entry") {
            int dir = 0, tc = 0;
            uint64_t rows = 0, tw = 0, ug = 0, xcd = 0, chain = 0, blk = 0, unr = 0;
            if (!(f >> dir >> tc >> rows >> tw >> ug >> xcd) || (version >= 2 && !(f >> chain)) ||
                (version >= 3 && !(f >> blk >> unr)) || (dir != GRGL_DIRECTION_UP && dir != GRGL_DIRECTION_DOWN) ||
                tc < 0 || tc > (int)kX8 || rows == 0 || rows > 65536 || tw > 64 || ug > 1 || xcd > 2 || chain > 3 ||
                (blk != 0 && blk != 256 && blk != 512 && blk != 1024) || (unr != 0 && unr != 4 && unr != 8))
                throw MisuseError("malformed tuning file (entry)");
            Tuning t;
            t.teamLanes = (uint32_t)tw;
            t.useGraph = (uint32_t)ug;
            t.xcdMap = (uint32_t)xcd;
            t.chain = (uint32_t)chain;
            t.blockThreads = (uint32_t)blk;
            t.loadsInFlight = (uint32_t)unr;
            entries.push_back({std::make_tuple(dir, tc, (uint32_t)rows), t});
        } else {
            throw MisuseError("malformed tuning file");
        }
    }
    std::lock_guard<std::mutex> lk(g->mu);
    for (const Dir& d : dirs) {
        DeviceDirPlan& plan = d.dir == GRGL_DIRECTION_UP ? g->up : g->down;
        if (plan.table->heavyFanin != (uint32_t)d.heavy || plan.table->chunkEntries != d.chunk)
            plan.table = plan.buildTable((uint32_t)d.heavy, d.chunk);
    }
    for (const auto& e : entries) g->tuning[e.first] = e.second;
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_kernel_timing(grgl_graph* g, int enable) {
    GRGL_API_BEGIN
    if (!g) throw MisuseError("null graph");
    g->timing = enable != 0;
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_kernel_timing_read(grgl_graph* g, double* total_ms, uint64_t* launches,
                                                    double* algorithmic_bytes, int reset) {
    GRGL_API_BEGIN
    if (!g) throw MisuseError("null graph");
    GRGL_HIP_CHECK(hipSetDevice(g->device));
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<double> bytes;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        events.swap(g->timingEvents);
        bytes.swap(g->timingBytes);
    }
    double ms = 0.0, moved = 0.0;
    hipError_t firstError = hipSuccess;
    for (size_t i = 0; i < events.size(); i++) {
        const hipError_t e = hipEventSynchronize(events[i].second);
        float one = 0.0f;
        if (e == hipSuccess) (void)hipEventElapsedTime(&one, events[i].first, events[i].second);
        else if (firstError == hipSuccess) firstError = e;
        ms += one;
        moved += bytes[i];
        (void)hipEventDestroy(events[i].first);
        (void)hipEventDestroy(events[i].second);
    }
    GRGL_HIP_CHECK(firstError);
    std::lock_guard<std::mutex> lk(g->mu);
    g->timedMs += ms;
    g->timedLaunches += events.size();
    g->timedBytes += moved;
    if (total_ms) *total_ms = g->timedMs;
    if (launches) *launches = g->timedLaunches;
    if (algorithmic_bytes) *algorithmic_bytes = g->timedBytes;
    if (reset) {
        g->timedMs = 0.0;
        g->timedLaunches = 0;
        g->timedBytes = 0.0;
    }
    GRGL_API_END
}

GRGL_EXPORT grgl_status grgl_hip_autotune(grgl_graph* g, grgl_direction direction, grgl_dtype dtype, uint32_t rows,
                                          grgl_tuning* chosen, float* best_ms) {
    GRGL_API_BEGIN
    if (!g) throw MisuseError("null graph");
    if (direction != GRGL_DIRECTION_UP && direction != GRGL_DIRECTION_DOWN) throw MisuseError("Invalid traversal direction");
    if (rows == 0 || rows > 65536) throw MisuseError("rows must be in 1..65536");
    const TypeClass tc = typeClassOf(dtype);
    const size_t es = dtypeSize(dtype);
    const RowGeometry geo = rowGeometry(rows, es);
    GRGL_HIP_CHECK(hipSetDevice(g->device));
    const bool savedTiming = g->timing.exchange(false);
    hipStream_t s = acquireInternalStream(g);
    Workspace* ws = acquireWorkspace(g, rowsBufferBytes(g, direction, rows, es), s);
    std::vector<uint32_t> tws;
    for (uint32_t tw = 1; tw <= maxTeamLanes(geo); tw *= 2) {
        if (levelTiles(geo, tw, es) <= 65535) tws.push_back(tw);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void* io = nullptr; // candidates' input and output, when timed as whole traversals
    Tuning best;
    float bestMs = 1e30f;
    // The stored choice, put back if timing fails (candidates are stored while timed).
    const auto tuningKey = std::make_tuple((int)direction, (int)tc, rows);
    bool hadTuning = false;
    Tuning previous;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        auto it = g->tuning.find(tuningKey);
        if (it != g->tuning.end()) {
            hadTuning = true;
            previous = it->second;
        }
    }
    try {
        GRGL_HIP_CHECK(hipEventCreate(&e0));
        GRGL_HIP_CHECK(hipEventCreate(&e1));
        GRGL_HIP_CHECK(hipMemsetAsync(ws->ptr, 0, rowsBufferBytes(g, direction, rows, es), s));
        // Candidates: every team width with direct or graph launches and both block orders
        // (the XCD-aware order only where it differs), for rows of at most 64 bytes the
        // chained single launch, and for rows of at most 512 bytes the grouped levels at every
        // team width (rows over 64 bytes as 64-byte slices; wider rows need more slices than
        // the slices save launches).
        std::vector<Tuning> cands;
        Tuning grouped;
        grouped.chain = 3;
        const bool groupable = geo.rowBytes <= 512 &&
                               groupFor(direction == GRGL_DIRECTION_UP ? g->up : g->down, *launchTable(g, direction),
                                        tc, geo, grouped, false, false) != nullptr;
        for (uint32_t tw : tws) {
            for (uint32_t useGraph = 0; useGraph < 2; useGraph++) {
                for (uint32_t xcd = 1; xcd <= 2; xcd++) {
                    if (xcd == 2 && levelTiles(geo, tw, es) % 8 != 0) continue; // same launch as xcd = 1
                    Tuning t;
                    t.teamLanes = tw;
                    t.useGraph = useGraph;
                    t.xcdMap = xcd;
                    t.chain = 1;
                    cands.push_back(t);
                }
            }
            if (geo.rowBytes <= 64) {
                Tuning t;
                t.teamLanes = tw;
                t.xcdMap = 1;
                t.chain = 2;
                cands.push_back(t);
            }
            // Grouped levels (the other levels at this team width), direct or graph launches.
            if (groupable) {
                for (uint32_t useGraph = 0; useGraph < 2; useGraph++) {
                    for (uint32_t blk : {256U, 1024U}) {
                        if (blk == 1024 && geo.rowBytes > 8) continue; // 256 threads only
                        Tuning t;
                        t.teamLanes = tw;
                        t.xcdMap = 1;
                        t.chain = 3;
                        t.useGraph = useGraph;
                        t.blockThreads = blk;
                        cands.push_back(t);
                    }
                }
            }
            // Wide-team shapes: 512-thread blocks and 8 loads in flight, direct launches.
            if (tileTunable(levelVb(geo, tw, es), tw) && tc != kW16 && tc != kW8 && tc != kX8) {
                for (uint32_t shape = 1; shape < 4; shape++) {
                    Tuning t;
                    t.teamLanes = tw;
                    t.xcdMap = 1;
                    t.chain = 1;
                    t.blockThreads = (shape & 1) ? 512 : 256;
                    t.loadsInFlight = (shape & 2) ? 8 : 4;
                    cands.push_back(t);
                }
            }
        }
        // Rows of at most 512 bytes: each candidate is timed as a whole traversal (seed, levels,
        // read-out), so that forms which also change the seed or the read-out (the folded seed
        // copy, the output written by the last level) and grouped levels are judged as calls
        // take them (DOWN K = 32 on the 200k x 1M graph: one launch per level chosen from the
        // level sequences alone, 0.419 ms a call; the grouped levels chosen from whole calls,
        // 0.387 ms).  Wider rows: the level sequence alone (their seeds and read-outs, 40% of a
        // 511-RHS call, are the same for every candidate and only add noise to the choice).
        const uint64_t inCols = direction == GRGL_DIRECTION_UP ? g->N : g->M;
        const uint64_t outCols = direction == GRGL_DIRECTION_UP ? g->M : g->N;
        const uint64_t ioBytes = (uint64_t)rows * (inCols + outCols) * es;
        if (geo.rowBytes <= 512 && ioBytes <= (4ULL << 30)) {
            GRGL_HIP_CHECK(hipMalloc(&io, ioBytes));
            GRGL_HIP_CHECK(hipMemsetAsync(io, 0, ioBytes, s));
        }
        auto runOnce = [&](const Tuning& t) {
            if (!io) {
                runLevels(g, direction, tc, geo, t, false, false, rows, es, ws->ptr, s, false, ws->chainState);
                return;
            }
            TravOptions opt;
            opt.chainState = ws->chainState;
            const std::vector<TravRequest> req{{rows, io, static_cast<char*>(io) + (size_t)rows * inCols * es, 0}};
            enqueueTraversal(g, dtype, direction, req, rows, opt, ws->ptr, s);
        };
        for (const Tuning& t : cands) {
            {
                std::lock_guard<std::mutex> lk(g->mu);
                g->tuning[tuningKey] = t; // the calls below take it (results never depend on it)
            }
            runOnce(t); // warm-up
            // Median of 3 trials; of 5 for level sequences alone (wide rows, where candidates
            // differ by a few percent and single trials by as much).
            const int nTrials = io ? 3 : 5;
            float trial[5];
            for (int k = 0; k < nTrials; k++) {
                GRGL_HIP_CHECK(hipEventRecord(e0, s));
                runOnce(t);
                GRGL_HIP_CHECK(hipEventRecord(e1, s));
                GRGL_HIP_CHECK(hipEventSynchronize(e1));
                GRGL_HIP_CHECK(hipEventElapsedTime(&trial[k], e0, e1));
            }
            std::sort(trial, trial + nTrials);
            if (trial[nTrials / 2] < bestMs) {
                bestMs = trial[nTrials / 2];
                best = t;
            }
        }
        checkChainHealth(g);
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (io) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(io);
        }
        {
            std::lock_guard<std::mutex> lk(g->mu);
            if (hadTuning)
                g->tuning[tuningKey] = previous;
            else
                g->tuning.erase(tuningKey);
        }
        releaseWorkspace(g, ws, s);
        releaseInternalStream(g, s);
        g->timing = savedTiming;
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (io) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(io);
    }
    releaseWorkspace(g, ws, s);
    releaseInternalStream(g, s);
    g->timing = savedTiming;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        g->tuning[std::make_tuple((int)direction, (int)tc, rows)] = best;
    }
    if (chosen) {
        chosen->team_lanes = best.teamLanes;
        const std::shared_ptr<const LaunchTable> tab = launchTable(g, direction);
        chosen->heavy_fanin = tab->heavyFanin;
        chosen->use_graph = best.useGraph;
        chosen->xcd_map = best.xcdMap;
        tileShape(tc, levelVb(geo, best.teamLanes, es), best.teamLanes, best, chosen->block_threads,
                  chosen->loads_in_flight);
        reportChain(direction == GRGL_DIRECTION_UP ? g->up : g->down, *tab, tc, geo, best, *chosen);
        chosen->max_launch_entries = (uint32_t)tab->chunkEntries;
    }
    if (best_ms) *best_ms = bestMs;
    GRGL_API_END
}
