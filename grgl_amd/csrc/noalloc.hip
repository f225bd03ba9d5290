// grgl_hip_matmul_noalloc: grgl_hip_matmul for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The traversal's row buffer, the only device memory a device-pointer grgl_hip_matmul call
// takes beyond its inputs and outputs, lives in `scratch`, a file-scope device array of
// GRGL_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the process ends.
// Its device address is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is handed to kernels: in host code `scratch` names a host
// stand-in, not device memory.
//
// Nothing in the row buffer has to start at zero: every row a traversal reads is written
// earlier in the same traversal (seeds, then level by level), and padding columns never reach
// an output.  So the contents a previous call left behind are never cleared.
//
// Every call on one GPU shares that GPU's array: calls that overlap in time on one GPU (from
// two threads, or on two streams without an order between them) race on it and corrupt each
// other's results.  Callers serialise them, e.g. on one stream.
#include <hip/hip_runtime.h>

#include <mutex>

#include "device_graph.h"

#define GRGL_EXPORT extern "C" __attribute__((visibility("default")))

static __device__ __align__(256) char scratch[GRGL_NOALLOC_SCRATCH_BYTES];

namespace grgl_amd {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
char* g_scratchPtr[kMaxDevices] = {};

// The device address of this device's `scratch` (the current device is `device`).
char* scratchAddress(int device) {
    if (device < 0 || device >= kMaxDevices) throw MisuseError("device ordinal out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        GRGL_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (p == nullptr) throw DeviceError("hipGetSymbolAddress returned no address for the scratch array");
        g_scratchPtr[device] = static_cast<char*>(p);
    }
    return g_scratchPtr[device];
}

} // namespace
} // namespace grgl_amd

using namespace grgl_amd;

GRGL_EXPORT grgl_status grgl_hip_matmul_noalloc(grgl_graph* g, grgl_dtype dtype, const void* input,
                                                size_t input_cols, size_t input_rows, grgl_direction direction,
                                                void* output, size_t output_size, int emit_all_nodes,
                                                int by_individual, const void* init, grgl_init_mode init_mode,
                                                void* miss, uint32_t flags, void* stream) {
    try {
        validateMatmulArgs(g, dtype, input, input_cols, input_rows, direction, output, output_size, emit_all_nodes,
                           by_individual, init, init_mode, miss);
        if (flags & GRGL_FLAG_HOST_PTRS)
            throw MisuseError("grgl_hip_matmul_noalloc takes device pointers only (GRGL_FLAG_HOST_PTRS needs staging "
                              "buffers)");
        const size_t es = dtypeSize(dtype);
        // Row blocks: the most input rows whose row buffer fits the scratch array (rows are
        // independent, so the result equals one traversal of all rows).
        size_t kb = std::min<size_t>(input_rows, 65536);
        while (kb > 0 && rowsBufferBytes(g, direction, (uint32_t)kb, es) > (size_t)GRGL_NOALLOC_SCRATCH_BYTES) kb /= 2;
        if (kb == 0)
            throw MisuseError("graph too large for grgl_hip_matmul_noalloc: one input row needs " +
                              std::to_string(rowsBufferBytes(g, direction, 1, es)) + " bytes of row buffer, the scratch "
                              "array holds " + std::to_string((size_t)GRGL_NOALLOC_SCRATCH_BYTES));
        GRGL_HIP_CHECK(hipSetDevice(g->device));
        char* rows = scratchAddress(g->device);
        hipStream_t s = (hipStream_t)stream;
        const size_t outCols = output_size / input_rows;
        TravOptions opt;
        opt.emitAll = emit_all_nodes != 0;
        opt.byIndividual = by_individual != 0;
        opt.initMode = init_mode;
        opt.accumulate = (flags & GRGL_FLAG_ACCUMULATE) != 0;
        opt.orderPreserving = (flags & GRGL_FLAG_ORDER_PRESERVING) != 0;
        opt.launchOnly = true;
        for (size_t k0 = 0; k0 < input_rows; k0 += kb) {
            const uint32_t K = (uint32_t)std::min<size_t>(kb, input_rows - k0);
            opt.init = init;
            if (init_mode == GRGL_INIT_VECTOR) opt.init = (const char*)init + k0 * es;
            else if (init_mode == GRGL_INIT_MATRIX) opt.init = (const char*)init + k0 * g->V * es;
            opt.miss = miss ? (char*)miss + k0 * g->M * es : nullptr;
            std::vector<TravRequest> reqs{{K, (const char*)input + k0 * input_cols * es, (char*)output + k0 * outCols * es,
                                           0}};
            enqueueTraversal(g, dtype, direction, reqs, K, opt, rows, s);
        }
    } catch (const std::exception& e) {
        return statusOf(e);
    }
    return GRGL_OK;
}
