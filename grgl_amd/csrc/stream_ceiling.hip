// Streaming bandwidth of the device (grgl_hip_stream_ceiling of grgl_hip.h): the read and copy
// rates a plain kernel reaches with one 16-byte access per thread of a 256-thread block, the
// access shape that runs fastest on MI355X.  The level kernels' HBM traffic is compared
// against these measured rates beside the published 8 TB/s.  And the random-gather rate
// (grgl_hip_gather_ceiling): 4-byte values gathered at random from a table, the access that
// bounds a traversal at one column, whose rows are 4 bytes.
#include <hip/hip_runtime.h>

#include <string>

#include "device_graph.h"

#define GRGL_EXPORT extern "C" __attribute__((visibility("default")))

namespace grgl_amd {
namespace {

constexpr int kProbeBlock = 256;

// XOR of every 16-byte word; the store happens only when the fold equals a value the probe's
// zero-filled buffer cannot give, so the loads stay live and nothing is written.
__global__ __launch_bounds__(kProbeBlock) void stream_read_kernel(const uint4* __restrict__ in, uint4* __restrict__ out,
                                                                  size_t n) {
    const size_t i = (size_t)blockIdx.x * kProbeBlock + threadIdx.x;
    if (i >= n) return;
    const uint4 v = in[i];
    if ((v.x ^ v.y ^ v.z ^ v.w) == 0x9e3779b9u) out[i] = v;
}

__global__ __launch_bounds__(kProbeBlock) void stream_copy_kernel(const uint4* __restrict__ in, uint4* __restrict__ out,
                                                                  size_t n) {
    const size_t i = (size_t)blockIdx.x * kProbeBlock + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// Random gathers of 4-byte values, the access of a level kernel at one column (K = 1): a row
// of kGatherRow consecutive list entries (coalesced index reads) whose values are gathered
// from a table at random positions, all loads issued before the adds.  One lane in `LPR`
// works, as in the level kernels' narrow teams (LPR = 4 leaves 3 lanes of 4 idle).
constexpr int kGatherRow = 16;

__global__ __launch_bounds__(kProbeBlock) void gather_fill_kernel(uint32_t* idx, size_t n, uint32_t tableLen) {
    const size_t i = (size_t)blockIdx.x * kProbeBlock + threadIdx.x;
    if (i >= n) return;
    uint64_t z = (uint64_t)i + 0x9e3779b97f4a7c15ULL; // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    idx[i] = (uint32_t)(z % tableLen);
}

template <int LPR>
__global__ __launch_bounds__(kProbeBlock) void gather_probe_kernel(const uint32_t* __restrict__ idx,
                                                                   const float* __restrict__ table,
                                                                   float* __restrict__ out, size_t rows) {
    const size_t t = (size_t)blockIdx.x * kProbeBlock + threadIdx.x;
    if (t % LPR != 0) return;
    const size_t r = t / LPR;
    if (r >= rows) return;
    const uint32_t* e = idx + r * kGatherRow;
    uint32_t j[kGatherRow];
#pragma unroll
    for (int u = 0; u < kGatherRow; u++) j[u] = e[u];
    float v[kGatherRow];
#pragma unroll
    for (int u = 0; u < kGatherRow; u++) v[u] = table[j[u]];
    float acc = 0.0f;
#pragma unroll
    for (int u = 0; u < kGatherRow; u++) acc += v[u];
    if (acc == 1234.5f) out[r] = acc; // keeps the loads live; the zeroed table never sums to it
}

struct ProbeBuffers {
    uint4* a = nullptr;
    uint4* b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~ProbeBuffers() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
    }
};

// Average milliseconds of `reps` launches after one untimed launch.
template <typename Launch> float timeLaunches(ProbeBuffers& p, int reps, Launch launch) {
    launch();
    GRGL_HIP_CHECK(hipGetLastError());
    GRGL_HIP_CHECK(hipDeviceSynchronize());
    GRGL_HIP_CHECK(hipEventRecord(p.e0, nullptr));
    for (int r = 0; r < reps; r++) launch();
    GRGL_HIP_CHECK(hipGetLastError());
    GRGL_HIP_CHECK(hipEventRecord(p.e1, nullptr));
    GRGL_HIP_CHECK(hipEventSynchronize(p.e1));
    float ms = 0.0f;
    GRGL_HIP_CHECK(hipEventElapsedTime(&ms, p.e0, p.e1));
    return ms / (float)reps;
}

} // namespace
} // namespace grgl_amd

using namespace grgl_amd;

GRGL_EXPORT grgl_status grgl_hip_gather_ceiling(int device, size_t table_bytes, uint32_t lanes_per_row,
                                                double* gathers_per_s) {
    try {
        if (!gathers_per_s) throw MisuseError("null argument");
        if (lanes_per_row != 1 && lanes_per_row != 4) throw MisuseError("lanes_per_row must be 1 or 4");
        const size_t tableLen = table_bytes / 4;
        if (tableLen < 1024 || tableLen >= (1ULL << 32)) throw MisuseError("gather_ceiling: table of 4 KB to 16 GB");
        int ndev = 0;
        GRGL_HIP_CHECK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw MisuseError("invalid device ordinal");
        GRGL_HIP_CHECK(hipSetDevice(device));
        const size_t rows = 1ULL << 20, n = rows * kGatherRow; // 16M gathers per launch
        ProbeBuffers p;
        GRGL_HIP_CHECK(hipMalloc(&p.a, n * 4));
        GRGL_HIP_CHECK(hipMalloc(&p.b, (tableLen * 4 + 15) / 16 * 16 + rows * 4));
        GRGL_HIP_CHECK(hipMemset(p.b, 0, (tableLen * 4 + 15) / 16 * 16 + rows * 4));
        GRGL_HIP_CHECK(hipEventCreate(&p.e0));
        GRGL_HIP_CHECK(hipEventCreate(&p.e1));
        uint32_t* idx = reinterpret_cast<uint32_t*>(p.a);
        float* table = reinterpret_cast<float*>(p.b);
        float* out = table + (tableLen + 3) / 4 * 4;
        gather_fill_kernel<<<dim3((uint32_t)((n + kProbeBlock - 1) / kProbeBlock)), kProbeBlock>>>(idx, n,
                                                                                                (uint32_t)tableLen);
        const dim3 grid((uint32_t)((rows * lanes_per_row + kProbeBlock - 1) / kProbeBlock));
        const float ms = timeLaunches(p, 10, [&] {
            if (lanes_per_row == 1) gather_probe_kernel<1><<<grid, kProbeBlock>>>(idx, table, out, rows);
            else gather_probe_kernel<4><<<grid, kProbeBlock>>>(idx, table, out, rows);
        });
        *gathers_per_s = (double)n / (ms * 1e-3);
        return GRGL_OK;
    } catch (const MisuseError& e) {
        setLastError(e.what());
        return GRGL_ERR_MISUSE;
    } catch (const NoMemError& e) {
        setLastError(e.what());
        return GRGL_ERR_NOMEM;
    } catch (const std::exception& e) {
        setLastError(e.what());
        return GRGL_ERR_DEVICE;
    }
}

GRGL_EXPORT grgl_status grgl_hip_stream_ceiling(int device, size_t bytes, double* read_gbps, double* copy_gbps) {
    try {
        if (!read_gbps || !copy_gbps) throw MisuseError("null argument");
        const size_t n = bytes / 16;
        if (n < kProbeBlock) throw MisuseError("stream_ceiling needs at least 4096 bytes");
        if (n / kProbeBlock >= (1ULL << 31) / kProbeBlock) throw MisuseError("stream_ceiling: buffer too large");
        int ndev = 0;
        GRGL_HIP_CHECK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw MisuseError("invalid device ordinal");
        GRGL_HIP_CHECK(hipSetDevice(device));
        ProbeBuffers p;
        GRGL_HIP_CHECK(hipMalloc(&p.a, n * 16));
        GRGL_HIP_CHECK(hipMalloc(&p.b, n * 16));
        GRGL_HIP_CHECK(hipMemset(p.a, 0, n * 16));
        GRGL_HIP_CHECK(hipMemset(p.b, 0, n * 16));
        GRGL_HIP_CHECK(hipEventCreate(&p.e0));
        GRGL_HIP_CHECK(hipEventCreate(&p.e1));
        const dim3 grid((uint32_t)((n + kProbeBlock - 1) / kProbeBlock));
        const float readMs = timeLaunches(p, 10, [&] { stream_read_kernel<<<grid, kProbeBlock>>>(p.a, p.b, n); });
        const float copyMs = timeLaunches(p, 10, [&] { stream_copy_kernel<<<grid, kProbeBlock>>>(p.a, p.b, n); });
        *read_gbps = (double)(n * 16) / (readMs * 1e6);
        *copy_gbps = 2.0 * (double)(n * 16) / (copyMs * 1e6);
        return GRGL_OK;
    } catch (const MisuseError& e) {
        setLastError(e.what());
        return GRGL_ERR_MISUSE;
    } catch (const NoMemError& e) {
        setLastError(e.what());
        return GRGL_ERR_NOMEM;
    } catch (const std::exception& e) {
        setLastError(e.what());
        return GRGL_ERR_DEVICE;
    }
}
