// Streaming bandwidth of the device (grgl_hip_stream_ceiling of grgl_hip.h): the read and copy
// rates a plain kernel reaches with one 16-byte access per thread of a 256-thread block, the
// access shape that runs fastest on MI355X.  The level kernels' HBM traffic is compared
// against these measured rates beside the published 8 TB/s.
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
