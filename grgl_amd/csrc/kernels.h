// gfx950 device kernels for GRG matrix multiplication.
//
// Data layout in HBM (one traversal):
//   rows      node-major value buffer: row r holds the K values of one node (or one seed
//             row), K contiguous, `ld` bytes apart.  Rows [0, numNodes) are the nodes in
//             level order; DOWN appends one row per mutation (input seeds) and, with
//             missing data, one per mutation for the missingness seeds.
//   off/src   per-row gather lists (u32 row indices), level by level.
//
// Kernels:
//   pull_level_kernel  one topological level: row j = [self +] sum of rows src[off[j]..].
//                      A "team" of TW lanes (TW*VB bytes of the row) owns a node and sums
//                      its list front to back.  Rows whose list is longer than the heavy
//                      threshold take a whole block: entry i goes to partial (i mod 64),
//                      each partial sums its entries in order, and the 64 partials meet in
//                      a fixed binary tree (strides 32,16,...,1).  Both schemes do the same
//                      arithmetic whatever TW or the K tiling is, so a request computed
//                      alone or stacked with others gives identical bits.
//   seed_rows_kernel   transposes a row-major K x C matrix into rows (optionally through a
//                      column map), adding into them or overwriting.
//   readout_kernel     transposes rows back into a row-major K x C output, summing `per`
//                      rows per output column in order (by-individual read-out).
//   init_rows_kernel   fills rows with 0, a per-column constant, or 2*coal counts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_graph.h"

namespace grgl_amd {

constexpr uint32_t kInvalidRow = 0x7fffffffU;
constexpr int kBlock = 256;
constexpr int kHeavyPartials = 64;

template <typename T, bool XOR> __device__ __forceinline__ T combine(T a, T b) {
    if constexpr (XOR) {
        return (T)(a ^ b);
    } else {
        return (T)(a + b);
    }
}

template <int VB> struct RawVec;
template <> struct RawVec<4> { using type = uint32_t; };
template <> struct RawVec<8> { using type = uint2; };
template <> struct RawVec<16> { using type = uint4; };
// A whole 32-byte row held by one lane (heavy rows of 32-byte rows, one wavefront per row):
// two 16-byte accesses.
struct Raw32 { uint4 a, b; };
template <> struct RawVec<32> { using type = Raw32; };
// A whole 64-byte row held by one lane (grouped levels): four 16-byte accesses.
struct Raw64 { uint4 a, b, c, d; };
template <> struct RawVec<64> { using type = Raw64; };

template <typename T, int VB> struct Pack {
    static constexpr int N = VB / (int)sizeof(T);
    T v[N];
};

template <typename T, int VB> __device__ __forceinline__ Pack<T, VB> load_pack(const char* p) {
    typename RawVec<VB>::type raw = *reinterpret_cast<const typename RawVec<VB>::type*>(p);
    Pack<T, VB> out;
    __builtin_memcpy(&out, &raw, VB);
    return out;
}

template <typename T, int VB> __device__ __forceinline__ void store_pack(char* p, const Pack<T, VB>& v) {
    typename RawVec<VB>::type raw;
    __builtin_memcpy(&raw, &v, VB);
    *reinterpret_cast<typename RawVec<VB>::type*>(p) = raw;
}

template <typename T, int VB> __device__ __forceinline__ Pack<T, VB> zero_pack() {
    Pack<T, VB> z;
#pragma unroll
    for (int i = 0; i < Pack<T, VB>::N; i++) z.v[i] = (T)0;
    return z;
}

// Lane-wise wrapping add of the 8- or 16-bit fields of two 32-bit words: the low bits of
// every field are added with the field's top bit cleared, so no carry crosses a field,
// and the top bit is restored by XOR (a + b mod 2^w per field).
template <typename T> __device__ __forceinline__ uint32_t swar_add(uint32_t x, uint32_t y) {
    constexpr uint32_t H = sizeof(T) == 1 ? 0x80808080U : 0x80008000U;
    return ((x & ~H) + (y & ~H)) ^ ((x ^ y) & H);
}

template <typename T, int VB, bool XOR>
__device__ __forceinline__ void acc_pack(Pack<T, VB>& a, const Pack<T, VB>& b) {
    if constexpr (sizeof(T) < 4) {
        uint32_t wa[VB / 4], wb[VB / 4];
        __builtin_memcpy(wa, &a, VB);
        __builtin_memcpy(wb, &b, VB);
#pragma unroll
        for (int i = 0; i < VB / 4; i++) wa[i] = XOR ? (wa[i] ^ wb[i]) : swar_add<T>(wa[i], wb[i]);
        __builtin_memcpy(&a, wa, VB);
    } else {
#pragma unroll
        for (int i = 0; i < Pack<T, VB>::N; i++) a.v[i] = combine<T, XOR>(a.v[i], b.v[i]);
    }
}

// Lane shuffle of a whole pack (32-bit words), from the lane `delta` above within the
// 64-lane wavefront.
template <typename T, int VB> __device__ __forceinline__ Pack<T, VB> shfl_down_pack(const Pack<T, VB>& v, unsigned delta) {
    static_assert(VB % 4 == 0, "packs are whole 32-bit words");
    int w[VB / 4];
    __builtin_memcpy(w, &v, VB);
#pragma unroll
    for (int i = 0; i < VB / 4; i++) w[i] = __shfl_down(w[i], delta, 64);
    Pack<T, VB> out;
    __builtin_memcpy(&out, w, VB);
    return out;
}

struct PullArgs {
    char* rows;
    const char* readRows; // rows the lists are read from: `rows`, or for a first level whose
                          // lists read only seed rows, the input itself (a one-column input is
                          // the seed rows' layout), based so that seed row r is at r * ld
    uint64_t ld;          // row stride, bytes
    const uint32_t* off;  // list offsets modulo 2^32
    const uint32_t* src;  // the launch's first gather entry
    uint32_t entryBase;   // off[] value (modulo 2^32) of that entry; lists are read at
                          // src[off[row] - entryBase], < 2^32 within a launch
    uint32_t levelBegin;  // first row of the launch (a level, or a chunk of a large level)
    uint32_t heavyCount;  // heavy rows at the front of the level
    uint32_t heavyBlocks; // blocks on heavy rows: heavyCount, or heavyCount / 4 rounded up
                          // when each wavefront takes a heavy row of its own (heavyPerWave)
    uint32_t heavyPerWave; // 1: rows of 4, 8, 16 bytes (32 bytes in the H32 variant); a
                           // wavefront's 64 lanes are a heavy row's 64 partials, each lane
                           // holding the whole row
    uint32_t levelEnd;
    uint32_t rowBytes;    // bytes of each row holding values (multiple of VB)
    uint32_t readSelf;    // 1: start from the row's current value (node initialisation)
    uint32_t xcdTiles;    // 0: 2-D grid (x = rows, y = column slices); else a 1-D grid over
                          // xcdTiles slices (a multiple of 8), see pull_block_coords
    // Output fusion (set only for the last DOWN level of a traversal of narrow rows, one
    // column slice, whose rows nothing reads afterwards): a row whose outCol entry c is below
    // outCols is written to column c of the K x outCols output (element k at k * outCols + c,
    // k < outK) instead of to its row, as 0 + value like the read-out it replaces; other rows
    // go to the row buffer.
    char* out = nullptr;
    const uint32_t* outCol = nullptr;
    uint32_t outCols = 0;
    uint32_t outK = 0;
    // Seed copy folded into a first level that reads its seeds from the input (readRows): the
    // first copyBlocks blocks (dispatched first, so the copy overlaps the level's rows) copy
    // copyBytes (a multiple of 4) from copySrc to copyDst, the seed rows the later levels read;
    // the level's own blocks follow.
    const char* copySrc = nullptr;
    char* copyDst = nullptr;
    uint64_t copyBytes = 0;
    uint32_t copyBlocks = 0;
};

// Stores row `row`'s finished value (its slice at colByte) where it goes: the row buffer, or
// with output fusion the output elements of its column.
template <typename T, int VB, bool XOR>
__device__ __forceinline__ void store_row(const PullArgs& a, uint32_t row, uint64_t colByte, const Pack<T, VB>& v) {
    if (a.outCol) {
        const uint32_t c = a.outCol[row];
        if (c < a.outCols) {
            Pack<T, VB> z = zero_pack<T, VB>();
            acc_pack<T, VB, XOR>(z, v);
            const uint32_t k0 = (uint32_t)(colByte / sizeof(T));
            T* out = reinterpret_cast<T*>(a.out);
#pragma unroll
            for (int e = 0; e < Pack<T, VB>::N; e++)
                if (k0 + (uint32_t)e < a.outK) out[(uint64_t)(k0 + (uint32_t)e) * a.outCols + c] = z.v[e];
            return;
        }
    }
    store_pack<T, VB>(a.rows + (uint64_t)row * a.ld + colByte, v);
}

// Workgroups are handed to the 8 XCDs round-robin by linear ID.  With a 1-D grid,
// block b runs on XCD b % 8; giving it column slice (b % 8) + 8 * ((b / 8) % (tiles / 8))
// makes each XCD work on its own slices, so each XCD's L2 holds 1/8 of the gathered bytes
// instead of every L2 caching the same slice.  It is an autotuning option, not the default:
// the 2-D grid runs the slices one after another, which keeps one slice of every row in the
// memory-side cache that all XCDs share, and measures faster on the 200k x 1M graph.
__device__ __forceinline__ void pull_block_coords(const PullArgs& a, uint32_t& bx, uint32_t& by) {
    if (a.xcdTiles == 0) {
        bx = blockIdx.x;
        by = blockIdx.y;
        return;
    }
    const uint32_t b = blockIdx.x, q = b >> 3, per = a.xcdTiles >> 3;
    by = (b & 7U) + 8U * (q % per);
    bx = q / per;
}

// Light rows: one team per row, list summed front to back with UNR row loads in flight.
// `idx(i)` returns gather entry i (from global memory or from an LDS copy).
// `b`, `e`: the row's list bounds relative to the launch's first entry.
template <typename T, int VB, int TW, bool XOR, int UNR, typename Idx>
__device__ __forceinline__ void pull_light_range(const PullArgs& a, uint32_t row, uint32_t b, uint32_t e,
                                                 uint64_t colByte, bool active, Idx idx) {
    using P = Pack<T, VB>;
    P acc = zero_pack<T, VB>();
    if (!active) return;
    if (a.readSelf) acc = load_pack<T, VB>(a.rows + (uint64_t)row * a.ld + colByte);
    // Batches of UNR entries: every load of a batch is issued before the first add, and
    // the adds run in list order.  The last, partial batch loads clamped (in-list) entries
    // unconditionally and adds only the real ones, so it is one round trip as well.
    uint32_t i = b;
    for (; i + UNR <= e; i += UNR) {
        P x[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) x[u] = load_pack<T, VB>(a.readRows + (uint64_t)idx(i + u) * a.ld + colByte);
#pragma unroll
        for (int u = 0; u < UNR; u++) acc_pack<T, VB, XOR>(acc, x[u]);
    }
    if (i < e) {
        P x[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const uint32_t j = i + u < e ? i + u : e - 1;
            x[u] = load_pack<T, VB>(a.readRows + (uint64_t)idx(j) * a.ld + colByte);
        }
#pragma unroll
        for (int u = 0; u < UNR; u++)
            if (i + u < e) acc_pack<T, VB, XOR>(acc, x[u]);
    }
    store_row<T, VB, XOR>(a, row, colByte, acc);
}

template <typename T, int VB, int TW, bool XOR, int UNR, typename Idx>
__device__ __forceinline__ void pull_light(const PullArgs& a, uint32_t row, uint64_t colByte, bool active,
                                           Idx idx) {
    pull_light_range<T, VB, TW, XOR, UNR>(a, row, a.off[row] - a.entryBase, a.off[row + 1] - a.entryBase, colByte,
                                          active, idx);
}

// Heavy rows: the whole block on one row, 64 interleaved partials, fixed reduction tree.
template <typename T, int VB, int TW, bool XOR, int BLK = kBlock>
__device__ __forceinline__ void pull_heavy(const PullArgs& a, uint32_t row, uint64_t colByte, bool active,
                                           uint32_t team, uint32_t lane) {
    using P = Pack<T, VB>;
    constexpr int NT = BLK / TW;                             // teams in the block
    constexpr int PPT = NT >= kHeavyPartials ? 1 : kHeavyPartials / NT; // partials per team
    constexpr int ACTIVE_TEAMS = NT >= kHeavyPartials ? kHeavyPartials : NT;
    __shared__ char lds[ACTIVE_TEAMS * TW * VB];
    const uint32_t b = a.off[row] - a.entryBase, e = a.off[row + 1] - a.entryBase;
    P part[PPT];
#pragma unroll
    for (int q = 0; q < PPT; q++) part[q] = zero_pack<T, VB>();
    constexpr int A = ACTIVE_TEAMS;                          // A * PPT == 64 always
    const bool working = active && team < (uint32_t)A;
    if (working) {
        // Entry b + i belongs to partial (i mod 64); team t owns partials t + q*A.
        // Walking i = t, t+A, t+2A, ... visits partial q = (step mod PPT), each partial's
        // entries in increasing order.
        // Each iteration covers STEPS steps (STEPS * PPT entries per team): all its loads
        // are issued before the adds, which run step by step, so each partial's order is
        // unchanged.  The last, partial iteration loads clamped (in-list) entries and adds
        // only the real ones.
        constexpr int STEPS = PPT >= 4 ? 1 : (VB * PPT <= 4 ? 16 : (VB * PPT <= 8 ? 8 : (VB <= 16 ? 4 / PPT : 4)));
        constexpr uint32_t STRIDE = (uint32_t)(STEPS * PPT) * A;
        uint32_t i = b + team;
        for (; i + STRIDE - A < e; i += STRIDE) {
            P x[STEPS][PPT];
#pragma unroll
            for (int st = 0; st < STEPS; st++)
#pragma unroll
                for (int q = 0; q < PPT; q++)
                    x[st][q] = load_pack<T, VB>(a.readRows + (uint64_t)a.src[i + (st * PPT + q) * A] * a.ld + colByte);
#pragma unroll
            for (int st = 0; st < STEPS; st++)
#pragma unroll
                for (int q = 0; q < PPT; q++) acc_pack<T, VB, XOR>(part[q], x[st][q]);
        }
        if (i < e) {
            P x[STEPS][PPT];
#pragma unroll
            for (int st = 0; st < STEPS; st++)
#pragma unroll
                for (int q = 0; q < PPT; q++) {
                    const uint32_t j = i + (uint32_t)(st * PPT + q) * A;
                    x[st][q] = load_pack<T, VB>(a.readRows + (uint64_t)a.src[j < e ? j : e - 1] * a.ld + colByte);
                }
#pragma unroll
            for (int st = 0; st < STEPS; st++)
#pragma unroll
                for (int q = 0; q < PPT; q++)
                    if (i + (uint32_t)(st * PPT + q) * A < e) acc_pack<T, VB, XOR>(part[q], x[st][q]);
        }
        // In-register tree steps: strides 32 .. A pair partials held by the same team
        // (partial t + q*A with partial t + (q + s/A)*A).
#pragma unroll
        for (int s = kHeavyPartials / 2; s >= A; s /= 2) {
#pragma unroll
            for (int q = 0; q < s / A; q++) acc_pack<T, VB, XOR>(part[q], part[q + s / A]);
        }
        if constexpr (TW == 1) {
            // Single-lane teams: the 64 partials are the 64 lanes of wavefront 0, so the
            // tree (strides 32 .. 1, partial t += partial t + s) runs on lane shuffles with
            // no LDS and no block barrier.  Every lane shuffles; lanes t < s combine.
            P mine = part[0];
#pragma unroll
            for (int s = kHeavyPartials / 2; s >= 1; s /= 2) {
                const P other = shfl_down_pack<T, VB>(mine, (unsigned)s);
                if (team < (uint32_t)s) acc_pack<T, VB, XOR>(mine, other);
            }
            if (team == 0) {
                if (a.readSelf) {
                    P self = load_pack<T, VB>(a.rows + (uint64_t)row * a.ld + colByte);
                    acc_pack<T, VB, XOR>(self, mine);
                    mine = self;
                }
                store_row<T, VB, XOR>(a, row, colByte, mine);
            }
        } else {
            store_pack<T, VB>(lds + ((uint64_t)team * TW + lane) * VB, part[0]);
        }
    }
    if constexpr (TW == 1) return;
    __syncthreads();
    // Remaining tree steps across teams through LDS: strides A/2 .. 1.
    for (int s = A / 2; s >= 1; s /= 2) {
        if (working && team < (uint32_t)s) {
            P mine = load_pack<T, VB>(lds + ((uint64_t)team * TW + lane) * VB);
            const P other = load_pack<T, VB>(lds + ((uint64_t)(team + s) * TW + lane) * VB);
            acc_pack<T, VB, XOR>(mine, other);
            store_pack<T, VB>(lds + ((uint64_t)team * TW + lane) * VB, mine);
        }
        __syncthreads();
    }
    if (working && team == 0) {
        P res = load_pack<T, VB>(lds + (uint64_t)lane * VB);
        if (a.readSelf) {
            P self = load_pack<T, VB>(a.rows + (uint64_t)row * a.ld + colByte);
            acc_pack<T, VB, XOR>(self, res);
            res = self;
        }
        store_row<T, VB, XOR>(a, row, colByte, res);
    }
}

// Narrow teams (TW <= 4, or accesses under 16 bytes: many rows per block, short row
// slices) are latency bound: the
// block's rows are consecutive, so their gather lists form one contiguous range of `src`,
// which is copied to LDS with coalesced loads (when it fits) before the teams walk their
// lists with 8 row loads in flight for 16-byte slices and 16 for narrower ones (16, or 32
// for 4- and 8-byte slices, for single-lane teams).  The summation order is the same in
// every variant.
constexpr uint32_t kStageEntries = 4096;

struct GlobalIdx {
    const uint32_t* src;
    __device__ uint32_t operator()(uint32_t i) const { return src[i]; }
};
struct StagedIdx {
    const uint32_t* lds;
    uint32_t base;
    __device__ uint32_t operator()(uint32_t i) const { return lds[i - base]; }
};

// Waiting hook of pull_level_block.  The level kernel needs none: kernel boundaries order the
// levels.  The chained kernel waits there for the previous level: poll() runs before a block
// barrier (thread 0 polls a completion counter), acquire() after it (false: the wait gave up,
// and the block does no work).
struct NoWait {
    static constexpr bool kWaits = false;
    __device__ void poll() const {}
    __device__ bool acquire() const { return true; }
};

// One block's share of a level: heavy rows (a block each, or 4 per block one per wavefront)
// or NT light rows.  The gather-list structure (offsets, indices) is read before the wait,
// the rows after it.  Returns false when the wait gave up (nothing written).
// BLK: threads per block; UNRW: row loads in flight per batch on the wide-team path (TW >= 8
// with 16-byte accesses).  Neither changes the arithmetic of any row.
template <typename T, int VB, int TW, bool XOR, bool H32, int BLK, int UNRW, typename Wait>
__device__ __forceinline__ bool pull_level_block(const PullArgs& a, uint32_t bx, uint32_t by, const Wait& wait) {
    const uint32_t lane = threadIdx.x % TW;
    const uint32_t team = threadIdx.x / TW;
    const uint64_t colByte = ((uint64_t)by * TW + lane) * VB;
    const bool active = colByte < a.rowBytes;
    constexpr uint32_t NT = BLK / TW;
    if (bx < a.heavyBlocks) {
        if constexpr (Wait::kWaits) {
            wait.poll();
            __syncthreads();
            if (!wait.acquire()) return false;
        }
        if (a.heavyPerWave) {
            // Short rows: one wavefront per heavy row, each lane one partial holding the
            // whole row (the single-lane-team arithmetic: same partials, same tree), so a
            // block takes 4 heavy rows and no lane idles.
            const uint32_t h = bx * (BLK / 64) + threadIdx.x / 64;
            if (h < a.heavyCount) {
                const uint32_t row = a.levelBegin + h, part = threadIdx.x % 64;
                if (a.rowBytes == 4) {
                    if constexpr (sizeof(T) <= 4) pull_heavy<T, 4, 1, XOR>(a, row, 0, true, part, 0);
                } else if (a.rowBytes == 8) {
                    if constexpr (sizeof(T) <= 8) pull_heavy<T, 8, 1, XOR>(a, row, 0, true, part, 0);
                } else if (a.rowBytes == 16) {
                    pull_heavy<T, 16, 1, XOR>(a, row, 0, true, part, 0);
                } else if constexpr (H32) {
                    pull_heavy<T, 32, 1, XOR>(a, row, 0, true, part, 0);
                }
            }
        } else {
            pull_heavy<T, VB, TW, XOR, BLK>(a, a.levelBegin + bx, colByte, active, team, lane);
        }
    } else if constexpr (TW <= 4 || VB < 16) {
        constexpr int UNR = TW == 1 ? (VB <= 8 ? 32 : 16) : (VB == 16 ? 8 : 16);
        __shared__ uint32_t staged[kStageEntries];
        const uint32_t r0 = a.levelBegin + a.heavyCount + (bx - a.heavyBlocks) * NT;
        const uint32_t r1 = r0 + NT < a.levelEnd ? r0 + NT : a.levelEnd;
        const uint32_t e0 = a.off[r0] - a.entryBase, n = a.off[r1] - a.off[r0];
        const uint32_t row = r0 + team;
        // The row's list bounds are loaded with the block's range, before the index copy, so
        // they cost no round trip of their own.
        uint32_t rb = 0, re = 0;
        if (row < r1) {
            rb = a.off[row] - a.entryBase;
            re = a.off[row + 1] - a.entryBase;
        }
        if (n <= kStageEntries) { // block-uniform
            // Every index load of a thread is issued before its first LDS store, so the
            // copy is one round trip whatever n is (a load-store loop costs one per pass).
            constexpr int SU = (int)(kStageEntries / BLK);
            uint32_t v[SU];
#pragma unroll
            for (int u = 0; u < SU; u++) {
                const uint32_t t = threadIdx.x + (uint32_t)u * BLK;
                if (t < n) v[u] = a.src[e0 + t];
            }
#pragma unroll
            for (int u = 0; u < SU; u++) {
                const uint32_t t = threadIdx.x + (uint32_t)u * BLK;
                if (t < n) staged[t] = v[u];
            }
            wait.poll();
            __syncthreads();
            if (!wait.acquire()) return false;
            if (row < r1) pull_light_range<T, VB, TW, XOR, UNR>(a, row, rb, re, colByte, active, StagedIdx{staged, e0});
        } else {
            if constexpr (Wait::kWaits) {
                wait.poll();
                __syncthreads();
                if (!wait.acquire()) return false;
            }
            if (row < r1) pull_light_range<T, VB, TW, XOR, UNR>(a, row, rb, re, colByte, active, GlobalIdx{a.src});
        }
    } else {
        if constexpr (Wait::kWaits) {
            wait.poll();
            __syncthreads();
            if (!wait.acquire()) return false;
        }
        const uint32_t row = a.levelBegin + a.heavyCount + (bx - a.heavyBlocks) * NT + team;
        if (row < a.levelEnd) pull_light<T, VB, TW, XOR, UNRW>(a, row, colByte, active, GlobalIdx{a.src});
    }
    return true;
}

// H32: the variant launched for levels with heavy rows when rows are 32 bytes; its
// per-wavefront heavy path holds 32 bytes per lane, whose registers would lower the
// occupancy of the levels without heavy rows, so they use the plain variant.
template <typename T, int VB, int TW, bool XOR, bool H32 = false, int BLK = kBlock, int UNRW = 4>
__global__ __launch_bounds__(BLK) void pull_level_kernel(PullArgs a) {
    if (blockIdx.x < a.copyBlocks) {
        // The folded seed copy (a one-column grid: blockIdx.x is the block's only coordinate).
        const uint64_t n = a.copyBytes / 4, stride = (uint64_t)a.copyBlocks * BLK;
        const uint32_t* src = reinterpret_cast<const uint32_t*>(a.copySrc);
        uint32_t* dst = reinterpret_cast<uint32_t*>(a.copyDst);
        for (uint64_t i = (uint64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += stride) dst[i] = src[i];
        return;
    }
    uint32_t bx, by;
    pull_block_coords(a, bx, by);
    bx -= a.copyBlocks;
    (void)pull_level_block<T, VB, TW, XOR, H32, BLK, UNRW>(a, bx, by, NoWait{});
}

// The trailing levels of a traversal that are too small to fill the chip (at most two
// passes of teams, a few KB of rows) run in ONE block, level after level, with a block
// barrier between levels instead of a kernel boundary.  Every row goes through the same
// pull_heavy / pull_light code as in pull_level_kernel, so results are bit-identical.
constexpr int kMaxTailLevels = 64;
struct TailArgs {
    char* rows;
    uint64_t ld;
    const uint32_t* off;
    const uint32_t* src;
    uint32_t entryBase;         // as PullArgs (src points at the first level's first entry)
    uint32_t rowBytes;
    uint32_t readSelf;
    uint32_t tiles;       // column slices of TW * VB bytes
    uint32_t numLevels;
    // Level boundaries and heavy-row counts travel in the kernel arguments (read with
    // uniform indices: scalar loads), so the kernel reads no table that retuning replaces.
    uint32_t levelStart[kMaxTailLevels + 1];
    uint32_t heavyCount[kMaxTailLevels]; // 0 for every level of an order-preserving traversal
};

template <typename T, int VB, int TW, bool XOR>
__global__ __launch_bounds__(kBlock) void pull_tail_kernel(TailArgs t) {
    const uint32_t lane = threadIdx.x % TW;
    const uint32_t team = threadIdx.x / TW;
    constexpr uint32_t NT = kBlock / TW;
    for (uint32_t l = 0; l < t.numLevels; l++) {
        PullArgs a;
        a.rows = t.rows;
        a.readRows = t.rows;
        a.ld = t.ld;
        a.off = t.off;
        a.src = t.src;
        a.entryBase = t.entryBase;
        a.levelBegin = t.levelStart[l];
        a.heavyCount = t.heavyCount[l];
        a.heavyBlocks = a.heavyCount;
        a.heavyPerWave = 0;
        a.levelEnd = t.levelStart[l + 1];
        a.rowBytes = t.rowBytes;
        a.readSelf = t.readSelf;
        for (uint32_t y = 0; y < t.tiles; y++) {
            const uint64_t colByte = ((uint64_t)y * TW + lane) * VB;
            const bool active = colByte < a.rowBytes;
            for (uint32_t h = 0; h < a.heavyCount; h++) // block-uniform; pull_heavy synchronises
                pull_heavy<T, VB, TW, XOR>(a, a.levelBegin + h, colByte, active, team, lane);
            for (uint32_t row = a.levelBegin + a.heavyCount + team; row < a.levelEnd; row += NT)
                pull_light<T, VB, TW, XOR, 4>(a, row, colByte, active, GlobalIdx{a.src});
        }
        __syncthreads(); // the next level reads rows written by other waves of this block
    }
}

// ---------------------------------------------------------------------------------------
// Grouped levels (grouping.h): block b runs its groups through levels [levelBegin, levelEnd)
// by itself, one segment (one level of its rows) after another with a block barrier in
// between; no block waits for another.  Each lane holds a whole row of RB bytes (rows of at
// most 64 bytes; wider rows as slices of 64 bytes, one per grid row).  The block keeps its
// rows' values in LDS slots.
//   Phase 0 loads everything the levels need, every load of the block independent of the
//   others: the row offsets of the block's slot lists, the slot lists, its segment table and
//   heavy rows' prefix lengths; the import slots (HBM rows that a list reads after its first
//   in-group entry); and each row's leading entries, which read HBM rows (earlier levels or
//   seed rows, complete before the launch), summed: a light row into its slot, starting from
//   the row's own value with node initialisation; a heavy row (longer than the heavy
//   threshold) into 64 partial slots, entry i to partial (i mod 64) as in pull_heavy.
//   Phase 1 continues each row's sum over its slot list, segment by segment, reading LDS
//   only: light rows one per lane; heavy rows one per wavefront, lane q finishing partial q,
//   then the fixed tree (strides 32..1) on lane shuffles and the row's own value added last,
//   as in pull_heavy.  So every row is the same sequence of adds as in the level kernels.
//   Rows are written to HBM once, after the last segment, so the barriers wait for no store
//   to memory.
// LDS of a block: value slots (its rows, 64 partials per heavy row, its imports) of RB bytes,
// 16-byte aligned, then 32-bit words: row offsets into the slot lists (rows + 1), the slot
// lists, 3 words per segment (first row, heavy rows, ordinal of its first heavy row) and
// each heavy row's prefix length (the layout groupLdsBytes counts).
struct GroupArgs {
    char* rows;
    uint64_t ld;
    const uint32_t* blockRow;    // blocks + 1: first group row of each block
    const uint32_t* blockSeg;    // blocks + 1: first segment of each block
    const uint32_t* blockHeavy;  // blocks + 1: first entry of heavyRow of each block
    const uint32_t* blockImport; // blocks + 1: first entry of importRow of each block
    const uint32_t* segRow;      // segments + 1: first group row of each segment
    const uint32_t* segHeavy;    // per segment: heavy rows at its front
    const uint32_t* segHeavyOrd; // per segment: block-relative ordinal of its first heavy row
    const uint32_t* row;         // per group row: the row it computes
    const uint32_t* off;         // group rows + 1: list bounds in src
    const uint32_t* pre;         // per group row: leading entries that read HBM rows
    const uint32_t* src;         // row indices (in-group ones tagged; only the prefix is read)
    const uint32_t* remOff;      // group rows + 1: slot list bounds in rem
    const uint32_t* rem;         // slot lists (block-relative slots)
    const uint32_t* heavyRow;    // heavy group rows of each block, in segment order
    const uint32_t* importRow;   // HBM row of each import slot, block after block
    char* scratch;               // heavy rows' 64 partial prefix sums, in heavyRow order
    uint32_t heavyFanin;
    uint32_t readSelf;
    uint32_t numRows, numHeavy;  // group rows and heavy rows of the plan (prefix launch)
    uint32_t lightBlocks;        // prefix launch: blocks of light rows before the heavy ones
    uint32_t prefixDone;         // 1: group_prefix_kernel ran before; 0: phase 0 sums the prefixes
};

// acc += rows[src[i]], rows[src[i + step]], ... below e, in that order, with UNR loads issued
// before the adds of each batch; the last, partial batch loads clamped (in-list) entries and
// adds only the real ones.
template <typename T, int RB, bool XOR, int UNR>
__device__ __forceinline__ void group_prefix_sum(const GroupArgs& a, Pack<T, RB>& acc, uint32_t i, uint32_t e,
                                                 uint32_t step) {
    using P = Pack<T, RB>;
    for (; i + (uint32_t)(UNR - 1) * step < e; i += (uint32_t)UNR * step) {
        P x[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) x[u] = load_pack<T, RB>(a.rows + (uint64_t)a.src[i + (uint32_t)u * step] * a.ld);
#pragma unroll
        for (int u = 0; u < UNR; u++) acc_pack<T, RB, XOR>(acc, x[u]);
    }
    if (i < e) {
        P x[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const uint32_t j = i + (uint32_t)u * step;
            x[u] = load_pack<T, RB>(a.rows + (uint64_t)a.src[j < e ? j : i] * a.ld);
        }
#pragma unroll
        for (int u = 0; u < UNR; u++)
            if (i + (uint32_t)u * step < e) acc_pack<T, RB, XOR>(acc, x[u]);
    }
}

// acc += slot[list[j]], slot[list[j + step]], ... below e, in that order (LDS only).
template <typename T, int RB, bool XOR>
__device__ __forceinline__ void group_slot_sum(const char* slots, const uint32_t* list, Pack<T, RB>& acc, uint32_t j,
                                               uint32_t e, uint32_t step) {
    using P = Pack<T, RB>;
    constexpr int U = 4;
    for (; j + (uint32_t)(U - 1) * step < e; j += (uint32_t)U * step) {
        P x[U];
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = load_pack<T, RB>(slots + (size_t)list[j + (uint32_t)u * step] * RB);
#pragma unroll
        for (int u = 0; u < U; u++) acc_pack<T, RB, XOR>(acc, x[u]);
    }
    for (; j < e; j += step) acc_pack<T, RB, XOR>(acc, load_pack<T, RB>(slots + (size_t)list[j] * RB));
}

// The leading HBM entries of every grouped row, summed over the whole device just before the
// grouped launch: one lane per light row, whose sum (from the row's own value with node
// initialisation) goes to the row's own place in the row buffer, which nothing reads before
// the grouped launch overwrites it; one wavefront per heavy row, whose 64 partial sums (entry i
// to partial i mod 64) go to `scratch`.  The adds are those pull_group_kernel would do.
// Light rows take one working lane in four (64 rows per block), as the level kernels' narrow
// teams do: a latency-bound pass over short lists runs faster spread over more wavefronts.
constexpr uint32_t kPrefixRowsPerBlock = kBlock / 4;

template <typename T, int RB, bool XOR>
__global__ __launch_bounds__(kBlock) void group_prefix_kernel(GroupArgs a) {
    using P = Pack<T, RB>;
    constexpr int UNR = RB <= 16 ? 8 : (RB == 32 ? 4 : 2);
    if (blockIdx.x < a.lightBlocks) {
        if (threadIdx.x % 4 != 0) return;
        const uint32_t i = blockIdx.x * kPrefixRowsPerBlock + threadIdx.x / 4;
        if (i >= a.numRows) return;
        const uint32_t lb = a.off[i], le = a.off[i + 1];
        if (le - lb > a.heavyFanin) return;
        char* self = a.rows + (uint64_t)a.row[i] * a.ld;
        P acc = a.readSelf ? load_pack<T, RB>(self) : zero_pack<T, RB>();
        group_prefix_sum<T, RB, XOR, UNR>(a, acc, lb, lb + a.pre[i], 1);
        store_pack<T, RB>(self, acc);
        return;
    }
    const uint32_t h = (blockIdx.x - a.lightBlocks) * (kBlock / 64) + threadIdx.x / 64, lane = threadIdx.x % 64;
    if (h >= a.numHeavy) return;
    const uint32_t i = a.heavyRow[h];
    const uint32_t lb = a.off[i];
    P acc = zero_pack<T, RB>();
    group_prefix_sum<T, RB, XOR, UNR>(a, acc, lb + lane, lb + a.pre[i], 64);
    store_pack<T, RB>(a.scratch + ((uint64_t)h * 64 + lane) * RB, acc);
}

template <typename T, int RB, bool XOR, int BLK>
__global__ __launch_bounds__(BLK) void pull_group_kernel(GroupArgs a) {
    extern __shared__ __align__(16) char slots[];
    using P = Pack<T, RB>;
    // Rows wider than RB bytes run as column slices of RB bytes, one per grid row: the columns
    // of a row never meet, so a slice is the same traversal over other bytes.
    a.rows += (uint64_t)blockIdx.y * RB;
    constexpr int UNR = RB <= 16 ? 8 : (RB == 32 ? 4 : 2);
    constexpr uint32_t kWaves = BLK / 64;
    const uint32_t b = blockIdx.x;
    const uint32_t r0 = a.blockRow[b], r1 = a.blockRow[b + 1];
    const uint32_t h0 = a.blockHeavy[b], h1 = a.blockHeavy[b + 1];
    const uint32_t s0 = a.blockSeg[b], s1 = a.blockSeg[b + 1];
    const uint32_t i0 = a.blockImport[b], i1 = a.blockImport[b + 1];
    const uint32_t nRows = r1 - r0, nHeavy = h1 - h0, nSeg = s1 - s0, nImport = i1 - i0;
    const uint32_t remBase = a.remOff[r0], nRem = a.remOff[r1] - remBase;
    const uint32_t importBase = nRows + 64 * nHeavy;
    const uint32_t valueBytes = ((importBase + nImport) * RB + 15) / 16 * 16;
    uint32_t* remOffL = reinterpret_cast<uint32_t*>(slots + valueBytes);
    uint32_t* remL = remOffL + nRows + 1;
    uint32_t* segL = remL + nRem;
    uint32_t* heavyPreL = segL + 3 * nSeg;
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    // Phase 0.
    for (uint32_t t = threadIdx.x; t <= nRows; t += BLK) remOffL[t] = a.remOff[r0 + t] - remBase;
    for (uint32_t t = threadIdx.x; t < nRem; t += BLK) remL[t] = a.rem[remBase + t];
    for (uint32_t t = threadIdx.x; t < nSeg; t += BLK) {
        segL[3 * t] = a.segRow[s0 + t] - r0;
        segL[3 * t + 1] = a.segHeavy[s0 + t];
        segL[3 * t + 2] = a.segHeavyOrd[s0 + t];
    }
    for (uint32_t t = threadIdx.x; t < nHeavy; t += BLK) heavyPreL[t] = a.pre[a.heavyRow[h0 + t]];
    for (uint32_t t = threadIdx.x; t < nImport; t += BLK)
        store_pack<T, RB>(slots + (size_t)(importBase + t) * RB,
                          load_pack<T, RB>(a.rows + (uint64_t)a.importRow[i0 + t] * a.ld));
    if (a.prefixDone) {
        // The prefix sums group_prefix_kernel left: a light row's in its own place (a heavy
        // row's slot is written only when the row completes, so copying its place too is
        // harmless), a heavy row's partials in the scratch area.  Four rows in flight per thread.
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += 4 * BLK) {
            P v[4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (i + (uint32_t)u * BLK < r1)
                    v[u] = load_pack<T, RB>(a.rows + (uint64_t)a.row[i + (uint32_t)u * BLK] * a.ld);
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (i + (uint32_t)u * BLK < r1) store_pack<T, RB>(slots + (size_t)(i + (uint32_t)u * BLK - r0) * RB, v[u]);
        }
        for (uint32_t k = wave; k < nHeavy; k += kWaves)
            store_pack<T, RB>(slots + (size_t)(nRows + 64 * k + lane) * RB,
                              load_pack<T, RB>(a.scratch + ((uint64_t)(h0 + k) * 64 + lane) * RB));
    } else {
        // Rows of at most 8 bytes: two rows per thread at a time, their list bounds, first four
        // prefix entries and values each issued for both rows before any add (DOWN K = 1-2 on
        // the 200k x 1M graph 2-8% faster; four rows raised the registers and lost occupancy).
        constexpr int RU = RB <= 8 ? 2 : 1;
        constexpr int FB = 4;
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += RU * BLK) {
            uint32_t lb[RU], n[RU];
            bool act[RU];
#pragma unroll
            for (int u = 0; u < RU; u++) {
                const uint32_t r = i + (uint32_t)u * BLK;
                act[u] = r < r1;
                lb[u] = act[u] ? a.off[r] : 0;
                const uint32_t le = act[u] ? a.off[r + 1] : 0;
                act[u] = act[u] && le - lb[u] <= a.heavyFanin;
                n[u] = act[u] ? a.pre[r] : 0;
            }
            P acc[RU];
#pragma unroll
            for (int u = 0; u < RU; u++)
                acc[u] = (act[u] && a.readSelf) ? load_pack<T, RB>(a.rows + (uint64_t)a.row[i + (uint32_t)u * BLK] * a.ld)
                                                : zero_pack<T, RB>();
            uint32_t e[RU][FB];
#pragma unroll
            for (int u = 0; u < RU; u++)
#pragma unroll
                for (int v = 0; v < FB; v++) e[u][v] = (uint32_t)v < n[u] ? a.src[lb[u] + (uint32_t)v] : 0;
            P x[RU][FB];
#pragma unroll
            for (int u = 0; u < RU; u++)
#pragma unroll
                for (int v = 0; v < FB; v++)
                    if ((uint32_t)v < n[u]) x[u][v] = load_pack<T, RB>(a.rows + (uint64_t)e[u][v] * a.ld);
#pragma unroll
            for (int u = 0; u < RU; u++) {
#pragma unroll
                for (int v = 0; v < FB; v++)
                    if ((uint32_t)v < n[u]) acc_pack<T, RB, XOR>(acc[u], x[u][v]);
                if (n[u] > (uint32_t)FB) group_prefix_sum<T, RB, XOR, UNR>(a, acc[u], lb[u] + FB, lb[u] + n[u], 1);
                if (act[u]) store_pack<T, RB>(slots + (size_t)(i + (uint32_t)u * BLK - r0) * RB, acc[u]);
            }
        }
        for (uint32_t k = wave; k < nHeavy; k += kWaves) {
            const uint32_t i = a.heavyRow[h0 + k];
            const uint32_t lb = a.off[i];
            P acc = zero_pack<T, RB>();
            group_prefix_sum<T, RB, XOR, UNR>(a, acc, lb + lane, lb + a.pre[i], 64);
            store_pack<T, RB>(slots + (size_t)(nRows + 64 * k + lane) * RB, acc);
        }
    }
    __syncthreads();
    // Phase 1.
    for (uint32_t s = 0; s < nSeg; s++) {
        const uint32_t g0 = segL[3 * s], g1 = s + 1 < nSeg ? segL[3 * (s + 1)] : nRows;
        const uint32_t hc = segL[3 * s + 1], hord = segL[3 * s + 2];
        for (uint32_t h = wave; h < hc; h += kWaves) { // wave-uniform
            const uint32_t i = g0 + h, k = hord + h;
            P acc = load_pack<T, RB>(slots + (size_t)(nRows + 64 * k + lane) * RB);
            // This partial's slot-list entries: list index pre + j with (pre + j) mod 64 == lane.
            const uint32_t j0 = (lane - heavyPreL[k]) & 63U;
            group_slot_sum<T, RB, XOR>(slots, remL, acc, remOffL[i] + j0, remOffL[i + 1], 64);
#pragma unroll
            for (int st = kHeavyPartials / 2; st >= 1; st /= 2) {
                const P other = shfl_down_pack<T, RB>(acc, (unsigned)st);
                if (lane < (uint32_t)st) acc_pack<T, RB, XOR>(acc, other);
            }
            if (lane == 0) {
                if (a.readSelf) {
                    P self = load_pack<T, RB>(a.rows + (uint64_t)a.row[r0 + i] * a.ld);
                    acc_pack<T, RB, XOR>(self, acc);
                    acc = self;
                }
                store_pack<T, RB>(slots + (size_t)i * RB, acc);
            }
        }
        for (uint32_t i = g0 + hc + threadIdx.x; i < g1; i += BLK) {
            P acc = load_pack<T, RB>(slots + (size_t)i * RB);
            group_slot_sum<T, RB, XOR>(slots, remL, acc, remOffL[i], remOffL[i + 1], 1);
            store_pack<T, RB>(slots + (size_t)i * RB, acc);
        }
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < nRows; i += BLK)
        store_pack<T, RB>(a.rows + (uint64_t)a.row[r0 + i] * a.ld, load_pack<T, RB>(slots + (size_t)i * RB));
}

// ---------------------------------------------------------------------------------------
// Chained traversal: every pulled level of a direction in ONE launch, with no grid barrier.
// The work is cut into the same per-block tasks as the level kernels (a level's heavy blocks,
// then its light blocks, levels in order) and numbered in that order.  Blocks claim task
// numbers from a ticket counter; a task of level l first reads its gather-list structure, then
// waits until every task of level l - 1 has completed (a per-level completion counter), then
// gathers and stores its rows and counts itself complete.
//   Progress: tickets are claimed in order and a block finishes its task before claiming the
//   next, so the lowest unfinished task is always held by a running block whose predecessors
//   are complete; no wait depends on a block that has not started, whatever the grid size or
//   residency.
//   Ordering: rows are stored, the block synchronises, and thread 0 increments the counter
//   with agent-scope release; the waiting block's thread 0 reads the counter, the block
//   synchronises, and every wavefront performs an agent-scope acquire before loading rows.
//   Each level's tasks waited for the level below, so a complete level l - 1 implies complete
//   levels below it (happens-before is transitive).
//   Bounded: a wait gives up after `waitTicks` of the constant 100 MHz clock, when another
//   block gave up (a flag in `state`), or when the host-mapped abort word is nonzero (read at
//   block start and on every 64th poll); a block that gives up stops claiming tasks, so the
//   kernel always ends.  The last block to leave sets the host-mapped `gaveUp` word if any
//   block gave up (a store: no atomics on host memory) and clears `state` (ticket, exits, flag,
//   counters) for the next traversal.
// Each row goes through pull_level_block, the level kernels' code, so results are bit for bit
// those of the level-by-level launches.
// ---------------------------------------------------------------------------------------
constexpr int kMaxChainLevels = 120;
constexpr uint32_t kChainHeader = 4; // state words: ticket, exits, gave-up flag, unused
static_assert((kChainHeader + kMaxChainLevels) * 4 <= kChainStateBytes, "chain state area");

struct ChainArgs {
    char* rows;
    uint64_t ld;
    const uint32_t* off;
    const uint32_t* src;        // all gather entries (fewer than 2^32)
    uint32_t rowBytes;
    uint32_t readSelf;
    uint32_t numLevels;         // pulled levels
    uint32_t totalTasks;
    uint32_t* state;            // device words, zero at launch: ticket, exits, flag, counters
    uint32_t* gaveUp;           // host-mapped: set to 1 when a traversal gave up
    const uint32_t* abortWord;  // host-mapped: nonzero = give up
    uint64_t waitTicks;         // longest wait, in ticks of wall_clock64()
    uint32_t taskStart[kMaxChainLevels + 1];  // first task of each level; [numLevels] = total
    uint32_t levelBegin[kMaxChainLevels + 1]; // first row of each level; [numLevels] = end
    uint32_t heavyCount[kMaxChainLevels];
    uint32_t heavyBlocks[kMaxChainLevels];    // bit 31: one heavy row per wavefront
};

struct ChainWait {
    static constexpr bool kWaits = true;
    const ChainArgs* c;
    uint32_t level;
    uint32_t* flag; // LDS: 1 = go, 0 = gave up
    __device__ bool giveUp() const {
        return __hip_atomic_load(c->state + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
               __hip_atomic_load(const_cast<uint32_t*>(c->abortWord), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0;
    }
    __device__ void poll() const {
        if (threadIdx.x != 0) return;
        uint32_t ok = 1;
        if (level > 0) {
            const uint32_t need = c->taskStart[level] - c->taskStart[level - 1];
            uint32_t* ctr = c->state + kChainHeader + level - 1;
            if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                const uint64_t t0 = wall_clock64();
                for (uint32_t spin = 0;; spin++) {
                    if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need) break;
                    if ((spin & 63) == 63 && (giveUp() || wall_clock64() - t0 > c->waitTicks)) {
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        *flag = ok;
    }
    __device__ bool acquire() const {
        if (*flag == 0) return false;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        return true;
    }
};

template <typename T, int VB, int TW, bool XOR, bool H32>
__global__ __launch_bounds__(kBlock) void pull_chain_kernel(ChainArgs c) {
    __shared__ uint32_t sTask, sOk;
    bool stop = false;
    if (threadIdx.x == 0 &&
        __hip_atomic_load(const_cast<uint32_t*>(c.abortWord), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0)
        __hip_atomic_store(c.state + 2, 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (!stop) {
        if (threadIdx.x == 0) {
            uint32_t t = c.totalTasks;
            if (__hip_atomic_load(c.state + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
                t = __hip_atomic_fetch_add(c.state, 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            sTask = t;
        }
        __syncthreads();
        const uint32_t task = sTask;
        if (task >= c.totalTasks) break; // block-uniform
        uint32_t lo = 0, hi = c.numLevels - 1; // last level whose first task is <= task
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) / 2;
            if (c.taskStart[mid] <= task) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t l = lo;
        PullArgs a;
        a.rows = c.rows;
        a.readRows = c.rows;
        a.ld = c.ld;
        a.off = c.off;
        a.src = c.src;
        a.entryBase = 0;
        a.levelBegin = c.levelBegin[l];
        a.heavyCount = c.heavyCount[l];
        a.heavyBlocks = c.heavyBlocks[l] & 0x7fffffffU;
        a.heavyPerWave = c.heavyBlocks[l] >> 31;
        a.levelEnd = c.levelBegin[l + 1];
        a.rowBytes = c.rowBytes;
        a.readSelf = c.readSelf;
        a.xcdTiles = 0;
        const bool done =
            pull_level_block<T, VB, TW, XOR, H32, kBlock, 4>(a, task - c.taskStart[l], 0, ChainWait{&c, l, &sOk});
        __syncthreads(); // the task's rows are stored; sTask and the LDS are free again
        if (threadIdx.x == 0) {
            if (done) __hip_atomic_fetch_add(c.state + kChainHeader + l, 1U, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            else __hip_atomic_store(c.state + 2, 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        stop = !done;
    }
    if (threadIdx.x == 0) {
        const uint32_t out = __hip_atomic_fetch_add(c.state + 1, 1U, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (out == gridDim.x - 1) {
            if (__hip_atomic_load(c.state + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
                __hip_atomic_store(c.gaveUp, 1U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            for (uint32_t i = 0; i < kChainHeader + c.numLevels; i++)
                __hip_atomic_store(c.state + i, 0U, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------------------
// Transposes between row-major K x C matrices and node rows, through an LDS tile of
// TK (k) x TC (columns) elements, TK = 2^tkShift <= 64 and TK * TC = kTileElems<T>.
// VR: elements per access on the node-row side (contiguous in k), VM: elements per
// access on the matrix side (contiguous in columns); 1 = scalar, 16/sizeof(T) = 16 B.
// The host picks the vector forms only when addresses, strides and offsets are 16-byte
// aligned.  A 16-byte row-side access may run past K into the row's padding (requests
// start 16-byte aligned, so that padding belongs to no other request); matrix-side
// accesses never cross the last column.
// ---------------------------------------------------------------------------------------
// Read-out tiles are 8 KB for 2- to 8-byte types (4 KB for bytes): two 16-byte accesses per
// thread and direction.  A plain copy on MI355X runs at 6.1 TB/s with one 16-byte access per
// thread of a 256-thread block, 5.8 with two and 5.6 with four (the stacked read-out of 511
// FP32 rows to 1M columns: 762 us with 16 KB tiles, 650 us with 8 KB).  Seed tiles stay at
// 16 KB for 4- and 8-byte types (the matching seed: 802 us with 16 KB, 809-838 with 8 KB).
// A tile must give every thread at least one 16-byte access (TE >= kBlock * 16 / sizeof(T)).
template <typename T> constexpr int kTileElems() { return sizeof(T) == 1 ? 4096 : 8192 / (int)sizeof(T); }
template <typename T> constexpr int kSeedTileElems() { return sizeof(T) == 8 ? 2048 : 4096; }
static_assert(kTileElems<uint8_t>() >= kBlock * 16 && kTileElems<uint16_t>() * 2 >= kBlock * 16 &&
                  kTileElems<float>() * 4 >= kBlock * 16 && kTileElems<double>() * 8 >= kBlock * 16 &&
                  kSeedTileElems<uint8_t>() >= kBlock * 16,
              "every thread needs at least one 16-byte access per tile");
constexpr uint32_t ilog2c(uint32_t v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }
// Tile index math uses shifts: TK, TC, VR and VM are powers of two, so every "vectors per
// tile line" count is one as well (a runtime integer division costs dozens of instructions).
template <typename T, int V> __device__ __forceinline__ uint32_t vecShift(uint32_t lineShift) {
    constexpr uint32_t vs = ilog2c((uint32_t)V);
    return lineShift > vs ? lineShift - vs : 0;
}

struct SeedArgs {
    const void* in;       // K x inCols, row-major
    uint64_t inCols;
    uint32_t K;
    char* rows;
    uint64_t ld;          // bytes
    uint32_t rowBase;
    uint32_t kOffset;     // first column of this request inside the row
    uint32_t D;           // destination rows
    uint32_t per;         // destination row d reads column d / per (no map)
    const uint32_t* colMap; // or column colMap[d]
    uint32_t accumulate;
    uint32_t tkShift;
    uint32_t normalizeBool;
};

template <typename T, bool XOR, int VR, int VM>
__global__ __launch_bounds__(kBlock) void seed_rows_kernel(SeedArgs a) {
    constexpr int TE = kSeedTileElems<T>();
    __shared__ T tile[TE + 64];
    const uint32_t TK = 1U << a.tkShift;
    const uint32_t TC = (uint32_t)TE >> a.tkShift;
    const uint64_t d0 = (uint64_t)blockIdx.x * TC;
    const uint32_t k0 = blockIdx.y * TK;
    const T* in = (const T*)a.in;
    const uint32_t tcShift = ilog2c((uint32_t)TE) - a.tkShift;
    // Load: k-major, VM columns per access; TE / VM accesses per tile, all issued before
    // the first LDS store so each thread keeps several global loads in flight.
    const uint32_t sK = vecShift<T, VM>(tcShift), vecPerK = 1U << sK;
    constexpr int NL = TE / VM / kBlock;
    T v[NL][VM];
#pragma unroll
    for (int it = 0; it < NL; it++) {
        const uint32_t e = threadIdx.x + it * kBlock;
        const uint32_t k = e >> sK, dv = (e & (vecPerK - 1)) * VM;
        const uint32_t gk = k0 + k;
        const uint64_t gd = d0 + dv;
#pragma unroll
        for (int i = 0; i < VM; i++) v[it][i] = (T)0;
        if (gk < a.K) {
            if constexpr (VM > 1) {
                if (gd + VM <= a.D) {
                    typename RawVec<16>::type raw = *reinterpret_cast<const typename RawVec<16>::type*>(
                        in + (uint64_t)gk * a.inCols + gd);
                    __builtin_memcpy(v[it], &raw, 16);
                } else {
                    for (int i = 0; i < VM; i++)
                        if (gd + i < a.D) v[it][i] = in[(uint64_t)gk * a.inCols + gd + i];
                }
            } else if (gd < a.D) {
                const uint64_t c = a.colMap ? (uint64_t)a.colMap[gd] : gd / a.per;
                v[it][0] = in[(uint64_t)gk * a.inCols + c];
            }
        }
    }
#pragma unroll
    for (int it = 0; it < NL; it++) {
        const uint32_t e = threadIdx.x + it * kBlock;
        const uint32_t k = e >> sK, dv = (e & (vecPerK - 1)) * VM;
#pragma unroll
        for (int i = 0; i < VM; i++) {
            T x = v[it][i];
            if (a.normalizeBool) x = (T)(x != (T)0);
            tile[k * (TC + 1) + dv + i] = x;
        }
    }
    __syncthreads();
    // Store: row-major over destination rows, VR k-values per access.
    // TE / VR accesses per tile (the host takes VR > 1 only when TK >= VR).
    const uint32_t sD = vecShift<T, VR>(a.tkShift), vecPerD = 1U << sD;
#pragma unroll
    for (int it = 0; it < TE / VR / kBlock; it++) {
        const uint32_t e = threadIdx.x + it * kBlock;
        const uint32_t d = e >> sD, k = (e & (vecPerD - 1)) * VR;
        const uint64_t gd = d0 + d;
        const uint32_t gk = k0 + k;
        if (gd >= a.D || gk >= a.K) continue;
        T* dst = (T*)(a.rows + (uint64_t)(a.rowBase + gd) * a.ld) + a.kOffset + gk;
        if constexpr (VR > 1) {
            T v[VR];
#pragma unroll
            for (int i = 0; i < VR; i++) v[i] = (k + i < TK) ? tile[(k + i) * (TC + 1) + d] : (T)0;
            if (a.accumulate) {
                T old[VR];
                typename RawVec<16>::type raw = *reinterpret_cast<const typename RawVec<16>::type*>(dst);
                __builtin_memcpy(old, &raw, 16);
#pragma unroll
                for (int i = 0; i < VR; i++) v[i] = combine<T, XOR>(old[i], v[i]);
            }
            typename RawVec<16>::type raw;
            __builtin_memcpy(&raw, v, 16);
            *reinterpret_cast<typename RawVec<16>::type*>(dst) = raw;
        } else {
            const T v = tile[k * (TC + 1) + d];
            *dst = a.accumulate ? combine<T, XOR>(*dst, v) : v;
        }
    }
}

struct ReadoutArgs {
    const char* rows;
    uint64_t ld;          // bytes
    uint32_t kOffset;
    uint32_t K;
    const uint32_t* map;  // C * per row indices (kInvalidRow = no contribution)
    uint32_t per;
    uint64_t C;
    void* out;            // K x outCols, row-major
    uint64_t outCols;
    uint32_t accumulate;
    uint32_t tkShift;
};

// Read-out of one row per output column (per == 1): batched row loads into the LDS tile,
// then each output vector is 0 (or the existing output) combined with the tile value and
// stored directly.  Unmapped columns hold 0 in the tile: without accumulation the result is
// 0 + 0 = 0; with accumulation they keep the existing output (nothing is added).
// `outRow(k)` points at output row k of the tile (column c0).
template <typename T, bool XOR, int VR, int VM, bool ACC, typename OutRow>
__device__ __forceinline__ void readout_single(const char* rows, uint64_t ld, uint32_t kOffset, uint32_t K,
                                               const uint32_t* map, uint64_t C, uint64_t c0,
                                               uint32_t k0, uint32_t TK, uint32_t TC, uint32_t sK, uint32_t vecPerK,
                                               uint32_t sC, uint32_t vecPerC, bool fullCols, OutRow outRow, T* tile,
                                               uint32_t* srcRow) {
    constexpr int TE = kTileElems<T>();
    for (uint32_t c = threadIdx.x; c < TC; c += kBlock) srcRow[c] = (c0 + c < C) ? map[c0 + c] : kInvalidRow;
    __syncthreads();
    constexpr int NL = TE / VR / kBlock;
    T v[NL][VR];
#pragma unroll
    for (int it = 0; it < NL; it++) {
        const uint32_t e = threadIdx.x + it * kBlock;
        const uint32_t c = e >> sC, k = (e & (vecPerC - 1)) * VR;
        const uint32_t gk = k0 + k;
        const uint32_t r = srcRow[c];
#pragma unroll
        for (int i = 0; i < VR; i++) v[it][i] = (T)0;
        if (r != kInvalidRow && gk < K) {
            const T* src = (const T*)(rows + (uint64_t)r * ld) + kOffset + gk;
            if constexpr (VR > 1) {
                typename RawVec<16>::type raw = *reinterpret_cast<const typename RawVec<16>::type*>(src);
                __builtin_memcpy(v[it], &raw, 16);
            } else {
                v[it][0] = *src;
            }
        }
    }
#pragma unroll
    for (int it = 0; it < NL; it++) {
        const uint32_t e = threadIdx.x + it * kBlock;
        const uint32_t c = e >> sC, k = (e & (vecPerC - 1)) * VR;
#pragma unroll
        for (int i = 0; i < VR; i++)
            if (k + i < TK) tile[(k + i) * (TC + 1) + c] = v[it][i];
    }
    __syncthreads();
    constexpr int OWN = TE / kBlock / VM;
#pragma unroll
    for (int o = 0; o < OWN; o++) {
        const uint32_t e = threadIdx.x + o * kBlock;
        const uint32_t k = e >> sK, c = (e & (vecPerK - 1)) * VM;
        if (k >= TK || k0 + k >= K) continue;
        T* dst = outRow(k);
        if (dst == nullptr) continue;
        dst += c;
        T w[VM];
        if constexpr (ACC) {
            if (VM > 1 && fullCols) {
                typename RawVec<16>::type raw = *reinterpret_cast<const typename RawVec<16>::type*>(dst);
                __builtin_memcpy(w, &raw, 16);
            } else {
                for (int i = 0; i < VM; i++) w[i] = (c0 + c + i < C) ? dst[i] : (T)0;
            }
#pragma unroll
            for (int i = 0; i < VM; i++)
                if (srcRow[c + i] != kInvalidRow) w[i] = combine<T, XOR>(w[i], tile[k * (TC + 1) + c + i]);
        } else {
#pragma unroll
            for (int i = 0; i < VM; i++) w[i] = combine<T, XOR>((T)0, tile[k * (TC + 1) + c + i]);
        }
        if constexpr (VM > 1) {
            if (fullCols) {
                typename RawVec<16>::type raw;
                __builtin_memcpy(&raw, w, 16);
                *reinterpret_cast<typename RawVec<16>::type*>(dst) = raw;
                continue;
            }
        }
        for (int i = 0; i < VM; i++)
            if (c0 + c + i < C) dst[i] = w[i];
    }
}

// The block's slice of the column map is staged in LDS once per summed row (`per`), so
// the load phase reads one row index per column and the sum phase checks validity in LDS.
// Unmapped columns contribute nothing: the written value is the initial value (0, or the
// existing output when accumulating), bit for bit.
// MODE: 0 = any `per` (rows summed per column), 1 = one row per column, 2 = one row per
// column added into the existing output.  The host picks the mode, so each instantiation
// carries only its own path (registers and branches of the others cost ~20% here).
template <typename T, bool XOR, int VR, int VM, int MODE>
__global__ __launch_bounds__(kBlock) void readout_kernel(ReadoutArgs a) {
    constexpr int TE = kTileElems<T>();
    constexpr int OWN = TE / kBlock / VM; // matrix-side vectors owned per thread
    __shared__ T tile[TE + 64];
    __shared__ uint32_t srcRow[TE];
    const uint32_t TK = 1U << a.tkShift;
    const uint32_t TC = (uint32_t)TE >> a.tkShift;
    const uint64_t c0 = (uint64_t)blockIdx.x * TC;
    const uint32_t k0 = blockIdx.y * TK;
    T* out = (T*)a.out;
    const uint32_t tcShift = ilog2c((uint32_t)TE) - a.tkShift;
    const uint32_t sK = vecShift<T, VM>(tcShift), vecPerK = 1U << sK;
    const bool fullCols = c0 + TC <= a.C;
    const uint32_t sC = vecShift<T, VR>(a.tkShift), vecPerC = 1U << sC;
    if constexpr (MODE != 0) {
        // One row per column: the sum is 0 (or the existing output) plus the row value,
        // written straight from the tile.
        readout_single<T, XOR, VR, VM, MODE == 2>(a.rows, a.ld, a.kOffset, a.K, a.map, a.C, c0, k0, TK, TC, sK,
                                                  vecPerK, sC, vecPerC, fullCols,
                                                  [&](uint32_t k) { return out + (uint64_t)(k0 + k) * a.outCols + c0; },
                                                  tile, srcRow);
        return;
    }
    T acc[OWN][VM];
#pragma unroll
    for (int o = 0; o < OWN; o++) {
        const uint32_t e = threadIdx.x + o * kBlock;
        const uint32_t k = e >> sK, c = (e & (vecPerK - 1)) * VM;
#pragma unroll
        for (int i = 0; i < VM; i++) acc[o][i] = (T)0;
        if (a.accumulate && k < TK && k0 + k < a.K) {
            const T* src = out + (uint64_t)(k0 + k) * a.outCols + c0 + c;
            if constexpr (VM > 1) {
                if (fullCols) {
                    typename RawVec<16>::type raw = *reinterpret_cast<const typename RawVec<16>::type*>(src);
                    __builtin_memcpy(acc[o], &raw, 16);
                    continue;
                }
            }
            for (int i = 0; i < VM; i++)
                if (c0 + c + i < a.C) acc[o][i] = src[i];
        }
    }
    for (uint32_t j = 0; j < a.per; j++) {
        for (uint32_t c = threadIdx.x; c < TC; c += kBlock)
            srcRow[c] = (c0 + c < a.C) ? a.map[(c0 + c) * a.per + j] : kInvalidRow;
        __syncthreads();
        // Load: one row slice per column, VR k-values per access; TE / VR accesses per tile
        // (the host takes VR > 1 only when TK >= VR), issued before the first LDS store.
        constexpr int NL = TE / VR / kBlock;
        T v[NL][VR];
#pragma unroll
        for (int it = 0; it < NL; it++) {
            const uint32_t e = threadIdx.x + it * kBlock;
            const uint32_t c = e >> sC, k = (e & (vecPerC - 1)) * VR;
            const uint32_t gk = k0 + k;
            const uint32_t r = srcRow[c];
#pragma unroll
            for (int i = 0; i < VR; i++) v[it][i] = (T)0;
            if (r != kInvalidRow && gk < a.K) {
                const T* src = (const T*)(a.rows + (uint64_t)r * a.ld) + a.kOffset + gk;
                if constexpr (VR > 1) {
                    typename RawVec<16>::type raw = *reinterpret_cast<const typename RawVec<16>::type*>(src);
                    __builtin_memcpy(v[it], &raw, 16);
                } else {
                    v[it][0] = *src;
                }
            }
        }
#pragma unroll
        for (int it = 0; it < NL; it++) {
            const uint32_t e = threadIdx.x + it * kBlock;
            const uint32_t c = e >> sC, k = (e & (vecPerC - 1)) * VR;
#pragma unroll
            for (int i = 0; i < VR; i++)
                if (k + i < TK) tile[(k + i) * (TC + 1) + c] = v[it][i];
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < OWN; o++) {
            const uint32_t e = threadIdx.x + o * kBlock;
            const uint32_t k = e >> sK, c = (e & (vecPerK - 1)) * VM;
            if (k < TK) {
#pragma unroll
                for (int i = 0; i < VM; i++)
                    if (srcRow[c + i] != kInvalidRow)
                        acc[o][i] = combine<T, XOR>(acc[o][i], tile[k * (TC + 1) + c + i]);
            }
        }
        if (j + 1 < a.per) __syncthreads(); // the next row of the sum reuses srcRow and the tile
    }
#pragma unroll
    for (int o = 0; o < OWN; o++) {
        const uint32_t e = threadIdx.x + o * kBlock;
        const uint32_t k = e >> sK, c = (e & (vecPerK - 1)) * VM;
        if (k >= TK || k0 + k >= a.K) continue;
        T* dst = out + (uint64_t)(k0 + k) * a.outCols + c0 + c;
        if constexpr (VM > 1) {
            if (fullCols) {
                typename RawVec<16>::type raw;
                __builtin_memcpy(&raw, acc[o], 16);
                *reinterpret_cast<typename RawVec<16>::type*>(dst) = raw;
                continue;
            }
        }
        for (int i = 0; i < VM; i++)
            if (c0 + c + i < a.C) dst[i] = acc[o][i];
    }
}

// ---------------------------------------------------------------------------------------
// Stacked transposes: the seeds (or read-outs) of every request of a batch in ONE launch.
// Column k of a stacked row belongs to the request whose [kOffset, kOffset + K) holds it
// (requests sorted by kOffset, in the kernel arguments); other columns are padding.  The
// seed writes whole stacked rows (padding as zero) with 16-byte vectors; the read-out reads
// whole row slices.  Per column the arithmetic is that of seed_rows_kernel /
// readout_kernel, so results are bit-identical to per-request launches.
// ---------------------------------------------------------------------------------------
constexpr int kMaxStack = 128;
struct StackReq {
    const void* ptr;  // request input (seed) or output (read-out), row-major K x cols
    uint32_t kOffset;
    uint32_t K;
};
struct StackArgs {
    char* rows;
    uint64_t ld;
    uint32_t rowBase;     // seed: first destination row
    uint32_t kCols;       // stacked columns to cover (row bytes / element size)
    uint64_t C;           // seed: destination rows; read-out: output columns
    uint64_t cols;        // columns of every request matrix
    uint32_t per;         // seed: row d reads column d / per; read-out: rows summed per column
    const uint32_t* map;  // read-out: C * per row indices
    uint32_t tkShift;
    uint32_t accumulate;  // read-out: add into the existing output
    uint32_t normalizeBool;
    uint32_t numReq;
    StackReq req[kMaxStack];
};

// kPtr[t] = first element of the request matrix row that holds stacked column k0 + t, or
// null for padding.  Resolved once per block into LDS.  The request table is read only with
// block-uniform indices (a uniform search for the first request that reaches the tile, then
// a walk over the requests overlapping it), so those reads are scalar loads through the
// scalar cache; per-lane indexing of the kernel-argument table would be vector loads from
// kernel-argument memory in every block.
__device__ __forceinline__ void stack_resolve(const StackArgs& a, uint32_t k0, uint32_t TK, char** kPtr, size_t es) {
    for (uint32_t t = threadIdx.x; t < TK; t += kBlock) kPtr[t] = nullptr;
    uint32_t lo = 0, hi = a.numReq; // first request with kOffset + K > k0
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (a.req[mid].kOffset + a.req[mid].K <= k0) lo = mid + 1;
        else hi = mid;
    }
    for (uint32_t r = lo; r < a.numReq && a.req[r].kOffset < k0 + TK; r++) {
        const uint32_t ko = a.req[r].kOffset, K = a.req[r].K;
        char* base = (char*)a.req[r].ptr;
        for (uint32_t t = threadIdx.x; t < TK; t += kBlock) {
            const uint32_t k = k0 + t;
            if (k >= ko && k < ko + K) kPtr[t] = base + (uint64_t)(k - ko) * a.cols * es;
        }
    }
}

template <typename T, bool XOR, int VR, int VM>
__global__ __launch_bounds__(kBlock) void seed_stacked_kernel(StackArgs a) {
    constexpr int TE = kSeedTileElems<T>();
    __shared__ T tile[TE + 64];
    __shared__ char* kPtr[64];
    const uint32_t TK = 1U << a.tkShift;
    const uint32_t TC = (uint32_t)TE >> a.tkShift;
    const uint64_t d0 = (uint64_t)blockIdx.x * TC;
    const uint32_t k0 = blockIdx.y * TK;
    stack_resolve(a, k0, TK, kPtr, sizeof(T));
    __syncthreads();
    const uint32_t tcShift = ilog2c((uint32_t)TE) - a.tkShift;
    const uint32_t sK = vecShift<T, VM>(tcShift), vecPerK = 1U << sK;
    constexpr int NL = TE / VM / kBlock;
    T v[NL][VM];
#pragma unroll
    for (int it = 0; it < NL; it++) {
        const uint32_t e = threadIdx.x + it * kBlock;
        const uint32_t k = e >> sK, dv = (e & (vecPerK - 1)) * VM;
        const uint64_t gd = d0 + dv;
        const T* in = (const T*)kPtr[k];
#pragma unroll
        for (int i = 0; i < VM; i++) v[it][i] = (T)0;
        if (in != nullptr) {
            if constexpr (VM > 1) {
                if (gd + VM <= a.C) {
                    typename RawVec<16>::type raw = *reinterpret_cast<const typename RawVec<16>::type*>(in + gd);
                    __builtin_memcpy(v[it], &raw, 16);
                } else {
                    for (int i = 0; i < VM; i++)
                        if (gd + i < a.C) v[it][i] = in[gd + i];
                }
            } else if (gd < a.C) {
                v[it][0] = in[gd / a.per];
            }
        }
    }
#pragma unroll
    for (int it = 0; it < NL; it++) {
        const uint32_t e = threadIdx.x + it * kBlock;
        const uint32_t k = e >> sK, dv = (e & (vecPerK - 1)) * VM;
#pragma unroll
        for (int i = 0; i < VM; i++) {
            T x = v[it][i];
            if (a.normalizeBool) x = (T)(x != (T)0);
            tile[k * (TC + 1) + dv + i] = x;
        }
    }
    __syncthreads();
    const uint32_t sD = vecShift<T, VR>(a.tkShift), vecPerD = 1U << sD;
#pragma unroll
    for (int it = 0; it < TE / VR / kBlock; it++) {
        const uint32_t e = threadIdx.x + it * kBlock;
        const uint32_t d = e >> sD, k = (e & (vecPerD - 1)) * VR;
        const uint64_t gd = d0 + d;
        const uint32_t gk = k0 + k;
        if (gd >= a.C || gk >= a.kCols) continue;
        T* dst = (T*)(a.rows + (uint64_t)(a.rowBase + gd) * a.ld) + gk;
        if constexpr (VR > 1) {
            T w[VR];
#pragma unroll
            for (int i = 0; i < VR; i++) w[i] = tile[(k + i) * (TC + 1) + d];
            typename RawVec<16>::type raw;
            __builtin_memcpy(&raw, w, 16);
            *reinterpret_cast<typename RawVec<16>::type*>(dst) = raw;
        } else {
            *dst = tile[k * (TC + 1) + d];
        }
    }
}

template <typename T, bool XOR, int VR, int VM, int MODE>
__global__ __launch_bounds__(kBlock) void readout_stacked_kernel(StackArgs a) {
    constexpr int TE = kTileElems<T>();
    constexpr int OWN = TE / kBlock / VM;
    __shared__ T tile[TE + 64];
    __shared__ uint32_t srcRow[TE];
    __shared__ char* kPtr[64];
    const uint32_t TK = 1U << a.tkShift;
    const uint32_t TC = (uint32_t)TE >> a.tkShift;
    const uint64_t c0 = (uint64_t)blockIdx.x * TC;
    const uint32_t k0 = blockIdx.y * TK;
    if constexpr (MODE != 0) {
        const uint32_t tcShift1 = ilog2c((uint32_t)TE) - a.tkShift;
        const uint32_t sK1 = vecShift<T, VM>(tcShift1), sC1 = vecShift<T, VR>(a.tkShift);
        stack_resolve(a, k0, TK, kPtr, sizeof(T)); // published by the helper's first barrier
        readout_single<T, XOR, VR, VM, MODE == 2>(a.rows, a.ld, 0, a.kCols, a.map, a.C, c0, k0, TK, TC, sK1,
                                                  1U << sK1, sC1, 1U << sC1, c0 + TC <= a.C,
                                                  [&](uint32_t k) { return kPtr[k] ? (T*)kPtr[k] + c0 : (T*)nullptr; },
                                                  tile, srcRow);
        return;
    }
    // The row pointers are needed by the accumulate preload and the final store only; without
    // accumulation they are resolved under the first map-slice barrier below.
    if (a.accumulate) {
        stack_resolve(a, k0, TK, kPtr, sizeof(T));
        __syncthreads();
    }
    const uint32_t tcShift = ilog2c((uint32_t)TE) - a.tkShift;
    const uint32_t sK = vecShift<T, VM>(tcShift), vecPerK = 1U << sK;
    const bool fullCols = c0 + TC <= a.C;
    T acc[OWN][VM];
#pragma unroll
    for (int o = 0; o < OWN; o++) {
        const uint32_t e = threadIdx.x + o * kBlock;
        const uint32_t k = e >> sK, c = (e & (vecPerK - 1)) * VM;
        const T* row = (const T*)kPtr[k];
#pragma unroll
        for (int i = 0; i < VM; i++) acc[o][i] = (T)0;
        if (a.accumulate && row != nullptr) {
            const T* src = row + c0 + c;
            if constexpr (VM > 1) {
                if (fullCols) {
                    typename RawVec<16>::type raw = *reinterpret_cast<const typename RawVec<16>::type*>(src);
                    __builtin_memcpy(acc[o], &raw, 16);
                    continue;
                }
            }
            for (int i = 0; i < VM; i++)
                if (c0 + c + i < a.C) acc[o][i] = src[i];
        }
    }
    const uint32_t sC = vecShift<T, VR>(a.tkShift), vecPerC = 1U << sC;
    for (uint32_t j = 0; j < a.per; j++) {
        for (uint32_t c = threadIdx.x; c < TC; c += kBlock)
            srcRow[c] = (c0 + c < a.C) ? a.map[(c0 + c) * a.per + j] : kInvalidRow;
        if (j == 0 && !a.accumulate) stack_resolve(a, k0, TK, kPtr, sizeof(T));
        __syncthreads();
        constexpr int NL = TE / VR / kBlock;
        T v[NL][VR];
#pragma unroll
        for (int it = 0; it < NL; it++) {
            const uint32_t e = threadIdx.x + it * kBlock;
            const uint32_t c = e >> sC, k = (e & (vecPerC - 1)) * VR;
            const uint32_t gk = k0 + k;
            const uint32_t r = srcRow[c];
#pragma unroll
            for (int i = 0; i < VR; i++) v[it][i] = (T)0;
            if (r != kInvalidRow && gk < a.kCols) {
                const T* src = (const T*)(a.rows + (uint64_t)r * a.ld) + gk;
                if constexpr (VR > 1) {
                    typename RawVec<16>::type raw = *reinterpret_cast<const typename RawVec<16>::type*>(src);
                    __builtin_memcpy(v[it], &raw, 16);
                } else {
                    v[it][0] = *src;
                }
            }
        }
#pragma unroll
        for (int it = 0; it < NL; it++) {
            const uint32_t e = threadIdx.x + it * kBlock;
            const uint32_t c = e >> sC, k = (e & (vecPerC - 1)) * VR;
#pragma unroll
            for (int i = 0; i < VR; i++)
                if (k + i < TK) tile[(k + i) * (TC + 1) + c] = v[it][i];
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < OWN; o++) {
            const uint32_t e = threadIdx.x + o * kBlock;
            const uint32_t k = e >> sK, c = (e & (vecPerK - 1)) * VM;
#pragma unroll
            for (int i = 0; i < VM; i++)
                if (srcRow[c + i] != kInvalidRow) acc[o][i] = combine<T, XOR>(acc[o][i], tile[k * (TC + 1) + c + i]);
        }
        if (j + 1 < a.per) __syncthreads(); // the next row of the sum reuses srcRow and the tile
    }
#pragma unroll
    for (int o = 0; o < OWN; o++) {
        const uint32_t e = threadIdx.x + o * kBlock;
        const uint32_t k = e >> sK, c = (e & (vecPerK - 1)) * VM;
        T* row = (T*)kPtr[k];
        if (row == nullptr) continue;
        T* dst = row + c0 + c;
        if constexpr (VM > 1) {
            if (fullCols) {
                typename RawVec<16>::type raw;
                __builtin_memcpy(&raw, acc[o], 16);
                *reinterpret_cast<typename RawVec<16>::type*>(dst) = raw;
                continue;
            }
        }
        for (int i = 0; i < VM; i++)
            if (c0 + c + i < a.C) dst[i] = acc[o][i];
    }
}

enum InitKind : uint32_t { kInitZero = 0, kInitVector = 1, kInitXtx = 3 };

struct InitArgs {
    char* rows;
    uint64_t ld;        // bytes
    uint32_t rowBegin;
    uint32_t rowEnd;
    uint32_t kOffset;
    uint32_t K;
    uint32_t kind;
    const void* vec;            // K values (kInitVector)
    const uint32_t* coalByRow;  // per row (kInitXtx)
};

template <typename T> __global__ __launch_bounds__(kBlock) void init_rows_kernel(InitArgs a) {
    const uint64_t n = (uint64_t)(a.rowEnd - a.rowBegin) * a.K;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const uint32_t r = a.rowBegin + (uint32_t)(i / a.K);
        const uint32_t k = (uint32_t)(i % a.K);
        T v = (T)0;
        if (a.kind == kInitVector) v = ((const T*)a.vec)[k];
        else if (a.kind == kInitXtx) v = (T)(a.coalByRow[r] * 2U);
        *((T*)(a.rows + (uint64_t)r * a.ld) + a.kOffset + k) = v;
    }
}

} // namespace grgl_amd
