// Resident-graph request loop with continuous batching (grgl_hip_server_* of grgl_hip.h).
//
// Requests (host buffers) queue per server.  One scheduler thread owns the device work:
// whenever a stream slot is free it takes the oldest queued request and stacks behind it
// every other queued request with the same (graph, direction, dtype), up to
// max_batch_rows rows, as long as the batch's device bytes fit the HBM budget left by
// the batches still running.  The batch is one traversal (grgl_amd::enqueueTraversal):
// stacked inputs are copied host -> pinned -> device, the level kernels run once for all
// of them, outputs come back through the slot's pinned buffer.  Batches on different
// slots overlap on the device.  A request's result does not depend on what it was
// batched with: every row of the traversal is computed by the same per-column
// arithmetic regardless of how many rows are stacked.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "device_graph.h"

#define GRGL_EXPORT extern "C" __attribute__((visibility("default")))

namespace grgl_amd {
namespace {

using Clock = std::chrono::steady_clock;

enum class ReqState { Queued, Running, Done, Cancelled, Failed };

struct ServerRequest {
    uint64_t id = 0;
    uint32_t graphId = 0;
    grgl_dtype dtype = GRGL_F32;
    grgl_direction dir = GRGL_DIRECTION_UP;
    uint32_t rows = 0;
    const void* hostIn = nullptr;
    void* hostOut = nullptr;
    bool device = false;          // in/out are device buffers used in place
    Clock::time_point submitted;
    ReqState state = ReqState::Queued;
    double latencyMs = 0.0;
    std::string error;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    char* raw = nullptr;            // device allocation: chain state area, then `dev`
    char* dev = nullptr;            // row buffer and staging of host requests
    uint32_t* chainState = nullptr; // kChainStateBytes, zero whenever no traversal runs
    size_t devBytes = 0;
    char* pinned = nullptr;
    size_t pinnedBytes = 0;
    bool busy = false;
    size_t budgetBytes = 0;
    grgl_graph* graph = nullptr;
    std::vector<std::shared_ptr<ServerRequest>> reqs;
    std::vector<uint32_t> offsets; // first stacked row of each request
    size_t inBytes = 0;   // per stacked row
    size_t outBytes = 0;  // per stacked row
    uint64_t edgeRhs = 0;
};

} // namespace

struct Server {
    grgl_server_config cfg{};
    std::vector<grgl_graph*> graphs;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<std::shared_ptr<ServerRequest>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<ServerRequest>> byId;
    uint64_t nextId = 1;
    uint64_t hbmInUse = 0, hbmPeak = 0;
    grgl_server_stats stats{};
    std::vector<double> latencies;
    bool stop = false;
    std::thread worker;
    // Batches completed on the device, counted by a stream host function so the scheduler
    // wakes as soon as a batch ends instead of polling.  The host function takes no lock and
    // makes no HIP call (the scheduler may hold `mu` inside a HIP call that waits for it).
    std::atomic<uint64_t> completions{0};

    static void onBatchDone(void* self) {
        Server* srv = static_cast<Server*>(self);
        srv->completions.fetch_add(1);
        srv->cvWork.notify_all();
    }
    void run();
    bool pollSlots();
    bool launchBatch(Slot& s, std::unique_lock<std::mutex>& lk);
};

namespace {

void growDevice(Slot& s, size_t bytes) {
    if (s.devBytes >= bytes) return;
    if (s.raw) GRGL_HIP_CHECK(hipFree(s.raw));
    s.raw = s.dev = nullptr;
    s.chainState = nullptr;
    s.devBytes = 0;
    const size_t sz = std::max<size_t>(bytes + bytes / 4, 8u << 20);
    GRGL_HIP_CHECK(hipMalloc(&s.raw, sz + kChainStateBytes));
    GRGL_HIP_CHECK(hipMemset(s.raw, 0, kChainStateBytes));
    s.chainState = reinterpret_cast<uint32_t*>(s.raw);
    s.dev = s.raw + kChainStateBytes;
    s.devBytes = sz;
}

void growPinned(Slot& s, size_t bytes) {
    if (s.pinnedBytes >= bytes) return;
    if (s.pinned) GRGL_HIP_CHECK(hipHostFree(s.pinned));
    s.pinned = nullptr;
    s.pinnedBytes = 0;
    const size_t sz = std::max<size_t>(bytes + bytes / 4, 8u << 20);
    GRGL_HIP_CHECK(hipHostMalloc(&s.pinned, sz, hipHostMallocDefault));
    s.pinnedBytes = sz;
}

size_t alignUp(size_t x) { return (x + 255) / 256 * 256; }

} // namespace

// Completes finished batches.  Called with mu held.
bool Server::pollSlots() {
    bool any = false;
    for (auto& s : slots) {
        if (!s.busy) continue;
        const hipError_t q = hipEventQuery(s.done);
        if (q == hipErrorNotReady) continue;
        const auto now = Clock::now();
        uint32_t Ktot = 0;
        for (size_t i = 0; i < s.reqs.size(); i++) Ktot = std::max(Ktot, s.offsets[i] + s.reqs[i]->rows);
        char* pinnedOut = s.pinned + alignUp(s.inBytes * Ktot);
        for (size_t i = 0; i < s.reqs.size(); i++) {
            auto& r = s.reqs[i];
            const size_t k0 = s.offsets[i];
            if (q == hipSuccess) {
                if (!r->device) std::memcpy(r->hostOut, pinnedOut + k0 * s.outBytes, (size_t)r->rows * s.outBytes);
                r->state = ReqState::Done;
            } else {
                r->state = ReqState::Failed;
                r->error = hipGetErrorString(q);
            }
            r->latencyMs = std::chrono::duration<double, std::milli>(now - r->submitted).count();
            latencies.push_back(r->latencyMs);
            stats.completed++;
        }
        stats.batches++;
        for (auto& r : s.reqs) stats.rows_processed += r->rows;
        stats.edge_rhs += s.edgeRhs;
        hbmInUse -= s.budgetBytes;
        s.budgetBytes = 0;
        s.reqs.clear();
        s.busy = false;
        any = true;
    }
    if (any) cvDone.notify_all();
    return any;
}

// Forms and launches one batch on slot s from the queue.  Called with mu held; the
// lock is kept (host copies are short and bounded by max_batch_rows).
bool Server::launchBatch(Slot& s, std::unique_lock<std::mutex>& lk) {
    (void)lk;
    // Drop cancelled requests at the head.
    while (!queue.empty() && queue.front()->state != ReqState::Queued) queue.pop_front();
    if (queue.empty()) return false;
    auto head = queue.front();
    grgl_graph* g = graphs[head->graphId];
    const size_t es = dtypeSize(head->dtype);
    const size_t inCols = head->dir == GRGL_DIRECTION_UP ? g->N : g->M;
    const size_t outCols = head->dir == GRGL_DIRECTION_UP ? g->M : g->N;
    // Device-buffer requests need no staging; the budget counts the row buffer plus staging
    // for the rows of host requests (charged as if every stacked row were staged).
    bool hostRows = false;
    auto batchBytes = [&](uint32_t K) {
        const size_t staging = hostRows ? alignUp(K * inCols * es) + alignUp(K * outCols * es) : 0;
        return alignUp(rowsBufferBytes(g, head->dir, K, es)) + staging;
    };
    std::vector<std::shared_ptr<ServerRequest>> picked;
    std::vector<uint32_t> offsets, rowsOf;
    uint32_t K = 0;
    const uint64_t budgetLeft = cfg.hbm_budget_bytes > hbmInUse ? cfg.hbm_budget_bytes - hbmInUse : 0;
    for (auto& r : queue) {
        if (r->state != ReqState::Queued) continue;
        if (r->graphId != head->graphId || r->dtype != head->dtype || r->dir != head->dir) continue;
        rowsOf.push_back(r->rows);
        const bool savedHost = hostRows;
        hostRows = hostRows || !r->device;
        uint64_t total = 0;
        std::vector<uint32_t> trial = stackOffsets(rowsOf, es, total);
        if ((!picked.empty() && total > cfg.max_batch_rows) || batchBytes((uint32_t)total) > budgetLeft) {
            rowsOf.pop_back();
            hostRows = savedHost;
            break;
        }
        picked.push_back(r);
        offsets = std::move(trial);
        K = (uint32_t)total;
    }
    if (picked.empty()) {
        // The head alone does not fit what is left of the budget.
        hostRows = !head->device;
        if (batchBytes(head->rows) > cfg.hbm_budget_bytes) {
            head->state = ReqState::Failed;
            head->error = "request exceeds the server's HBM budget";
            queue.pop_front();
            stats.rejected++;
            cvDone.notify_all();
            return true;
        }
        return false; // wait for running batches to release budget
    }
    const size_t bytes = batchBytes(K);
    const size_t rowsBytes = alignUp(rowsBufferBytes(g, head->dir, K, es));
    try {
        GRGL_HIP_CHECK(hipSetDevice(cfg.device));
        growDevice(s, bytes);
        if (hostRows) growPinned(s, alignUp(K * inCols * es) + alignUp(K * outCols * es));
    } catch (const std::exception& e) {
        for (auto& r : picked) {
            r->state = ReqState::Failed;
            r->error = e.what();
        }
        queue.erase(std::remove_if(queue.begin(), queue.end(),
                                   [](const std::shared_ptr<ServerRequest>& r) { return r->state != ReqState::Queued; }),
                    queue.end());
        cvDone.notify_all();
        return true;
    }
    for (auto& r : picked) r->state = ReqState::Running;
    queue.erase(std::remove_if(queue.begin(), queue.end(),
                               [](const std::shared_ptr<ServerRequest>& r) { return r->state != ReqState::Queued; }),
                queue.end());
    s.busy = true;
    s.graph = g;
    s.reqs = picked;
    s.inBytes = inCols * es;
    s.outBytes = outCols * es;
    s.budgetBytes = bytes;
    const DeviceDirPlan& plan = head->dir == GRGL_DIRECTION_UP ? g->up : g->down;
    s.edgeRhs = plan.numEntries * (uint64_t)K;
    hbmInUse += bytes;
    hbmPeak = std::max(hbmPeak, hbmInUse);
    char* devIn = s.dev + rowsBytes;
    char* devOut = devIn + alignUp(K * inCols * es);
    char* pinIn = s.pinned;
    std::vector<TravRequest> treqs;
    bool anyHost = false;
    for (size_t i = 0; i < picked.size(); i++) {
        auto& r = picked[i];
        const uint32_t k0 = offsets[i];
        if (r->device) {
            treqs.push_back({r->rows, r->hostIn, r->hostOut, k0});
        } else {
            anyHost = true;
            std::memcpy(pinIn + (size_t)k0 * inCols * es, r->hostIn, (size_t)r->rows * inCols * es);
            treqs.push_back({r->rows, devIn + (size_t)k0 * inCols * es, devOut + (size_t)k0 * outCols * es, k0});
        }
    }
    s.offsets = offsets;
    try {
        if (anyHost)
            GRGL_HIP_CHECK(hipMemcpyAsync(devIn, pinIn, (size_t)K * inCols * es, hipMemcpyHostToDevice, s.stream));
        TravOptions opt;
        opt.orderPreserving = (cfg.flags & GRGL_FLAG_ORDER_PRESERVING) != 0;
        opt.chainState = s.chainState;
        enqueueTraversal(g, head->dtype, head->dir, treqs, K, opt, s.dev, s.stream);
        if (anyHost)
            GRGL_HIP_CHECK(hipMemcpyAsync(s.pinned + alignUp(K * inCols * es), devOut, (size_t)K * outCols * es,
                                          hipMemcpyDeviceToHost, s.stream));
        GRGL_HIP_CHECK(hipEventRecord(s.done, s.stream));
        GRGL_HIP_CHECK(hipLaunchHostFunc(s.stream, &Server::onBatchDone, this));
    } catch (const std::exception& e) {
        // Wait for whatever was enqueued, then fail the batch.
        (void)hipStreamSynchronize(s.stream);
        for (auto& r : picked) {
            r->state = ReqState::Failed;
            r->error = e.what();
        }
        hbmInUse -= bytes;
        s.budgetBytes = 0;
        s.reqs.clear();
        s.busy = false;
        cvDone.notify_all();
    }
    return true;
}

void Server::run() {
    (void)hipSetDevice(cfg.device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        const uint64_t seenCompletions = completions.load();
        bool progressed = pollSlots();
        if (stop) {
            bool busy = false;
            for (auto& s : slots) busy |= s.busy;
            if (!busy) break;
        }
        // Batching window: hold a free slot until the oldest queued request has waited
        // batch_window_us, so later arrivals can share its traversal.
        bool windowOpen = false;
        Clock::time_point windowEnd;
        if (!queue.empty() && cfg.batch_window_us > 0 && !stop) {
            windowEnd = queue.front()->submitted + std::chrono::microseconds(cfg.batch_window_us);
            windowOpen = Clock::now() < windowEnd;
        }
        for (auto& s : slots) {
            if (s.busy || stop || windowOpen) continue;
            if (queue.empty()) break;
            progressed |= launchBatch(s, lk);
        }
        bool busy = false;
        for (auto& s : slots) busy |= s.busy;
        if (!progressed) {
            if (windowOpen) {
                cvWork.wait_until(lk, windowEnd);
            } else if (busy) {
                // Batches in flight: sleep until one completes (onBatchDone) or new work
                // arrives; the timeout only bounds a wake-up lost between the check and the wait.
                const size_t seenQueue = queue.size();
                cvWork.wait_for(lk, std::chrono::microseconds(500), [&] {
                    return stop || completions.load() != seenCompletions || queue.size() != seenQueue;
                });
            } else {
                cvWork.wait(lk, [&] { return stop || !queue.empty(); });
            }
        }
    }
}

} // namespace grgl_amd

using namespace grgl_amd;

namespace {

grgl_status fail(grgl_status st, const std::string& msg) {
    setLastError(msg);
    return st;
}

Server* S(grgl_server* s) { return reinterpret_cast<Server*>(s); }

grgl_status finish(Server* srv, uint64_t id, double* latencyMs, bool wait) {
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->byId.find(id);
    if (it == srv->byId.end()) return fail(GRGL_ERR_MISUSE, "unknown request id");
    auto r = it->second;
    if (wait) {
        srv->cvDone.wait(lk, [&] { return r->state == ReqState::Done || r->state == ReqState::Failed ||
                                          r->state == ReqState::Cancelled; });
    }
    switch (r->state) {
    case ReqState::Queued:
    case ReqState::Running: return GRGL_ERR_NOT_READY;
    case ReqState::Cancelled: srv->byId.erase(it); return fail(GRGL_ERR_CANCELLED, "request was cancelled");
    case ReqState::Failed: {
        const std::string e = r->error;
        srv->byId.erase(it);
        return fail(e.find("HBM budget") != std::string::npos ? GRGL_ERR_NOMEM : GRGL_ERR_DEVICE, e);
    }
    case ReqState::Done:
        if (latencyMs) *latencyMs = r->latencyMs;
        srv->byId.erase(it);
        return GRGL_OK;
    }
    return GRGL_OK;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    const size_t k = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
    std::nth_element(v.begin(), v.begin() + k, v.end());
    return v[k];
}

} // namespace

GRGL_EXPORT grgl_status grgl_hip_server_create(const grgl_server_config* config, grgl_server** out) {
    try {
        if (!out) return fail(GRGL_ERR_MISUSE, "null argument");
        *out = nullptr;
        std::unique_ptr<Server> srv(new Server());
        if (config) srv->cfg = *config;
        if (srv->cfg.num_streams == 0) srv->cfg.num_streams = 2;
        if (srv->cfg.max_batch_rows == 0) srv->cfg.max_batch_rows = 512;
        if (srv->cfg.max_batch_rows > 65536) return fail(GRGL_ERR_MISUSE, "max_batch_rows must be at most 65536");
        if (srv->cfg.hbm_budget_bytes == 0) srv->cfg.hbm_budget_bytes = 8ULL << 30;
        if (srv->cfg.max_queue == 0) srv->cfg.max_queue = 4096;
        int ndev = 0;
        GRGL_HIP_CHECK(hipGetDeviceCount(&ndev));
        if (srv->cfg.device < 0 || srv->cfg.device >= ndev) return fail(GRGL_ERR_MISUSE, "invalid device ordinal");
        GRGL_HIP_CHECK(hipSetDevice(srv->cfg.device));
        srv->slots.resize(srv->cfg.num_streams);
        for (auto& s : srv->slots) {
            GRGL_HIP_CHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
            GRGL_HIP_CHECK(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
        }
        Server* raw = srv.get();
        srv->worker = std::thread([raw] { raw->run(); });
        *out = reinterpret_cast<grgl_server*>(srv.release());
        return GRGL_OK;
    } catch (const std::exception& e) {
        return fail(GRGL_ERR_DEVICE, e.what());
    }
}

GRGL_EXPORT grgl_status grgl_hip_server_add_graph(grgl_server* server, grgl_graph* graph, uint32_t* graph_id) {
    if (!server || !graph || !graph_id) return fail(GRGL_ERR_MISUSE, "null argument");
    Server* srv = S(server);
    if (graph->device != srv->cfg.device) return fail(GRGL_ERR_MISUSE, "graph lives on another device");
    std::lock_guard<std::mutex> lk(srv->mu);
    *graph_id = (uint32_t)srv->graphs.size();
    srv->graphs.push_back(graph);
    return GRGL_OK;
}

namespace {
grgl_status submitImpl(grgl_server* server, uint32_t graph_id, grgl_dtype dtype, grgl_direction direction,
                       uint32_t rows, const void* host_input, void* host_output, bool device, uint64_t* request_id) {
    if (!server || !request_id || !host_input || !host_output) return fail(GRGL_ERR_MISUSE, "null argument");
    Server* srv = S(server);
    try {
        (void)typeClassOf(dtype);
    } catch (const std::exception& e) {
        return fail(GRGL_ERR_MISUSE, e.what());
    }
    if (direction != GRGL_DIRECTION_UP && direction != GRGL_DIRECTION_DOWN)
        return fail(GRGL_ERR_MISUSE, "Invalid traversal direction");
    if (rows == 0) return fail(GRGL_ERR_MISUSE, "matmul() requires non-zero dimensions.");
    if (rows > srv->cfg.max_batch_rows) return fail(GRGL_ERR_MISUSE, "request has more rows than max_batch_rows");
    auto r = std::make_shared<ServerRequest>();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (graph_id >= srv->graphs.size()) return fail(GRGL_ERR_MISUSE, "unknown graph id");
        srv->stats.submitted++;
        if (srv->queue.size() >= srv->cfg.max_queue) {
            srv->stats.rejected++;
            return fail(GRGL_ERR_BUSY, "server queue is full");
        }
        r->id = srv->nextId++;
        r->graphId = graph_id;
        r->dtype = dtype;
        r->dir = direction;
        r->rows = rows;
        r->hostIn = host_input;
        r->hostOut = host_output;
        r->device = device;
        r->submitted = Clock::now();
        srv->queue.push_back(r);
        srv->byId[r->id] = r;
        *request_id = r->id;
    }
    srv->cvWork.notify_one();
    return GRGL_OK;
}
} // namespace

GRGL_EXPORT grgl_status grgl_hip_server_submit(grgl_server* server, uint32_t graph_id, grgl_dtype dtype,
                                               grgl_direction direction, uint32_t rows, const void* host_input,
                                               void* host_output, uint64_t* request_id) {
    return submitImpl(server, graph_id, dtype, direction, rows, host_input, host_output, false, request_id);
}

GRGL_EXPORT grgl_status grgl_hip_server_submit_device(grgl_server* server, uint32_t graph_id, grgl_dtype dtype,
                                                      grgl_direction direction, uint32_t rows,
                                                      const void* device_input, void* device_output,
                                                      uint64_t* request_id) {
    return submitImpl(server, graph_id, dtype, direction, rows, device_input, device_output, true, request_id);
}

GRGL_EXPORT grgl_status grgl_hip_server_poll(grgl_server* server, uint64_t request_id, double* latency_ms) {
    if (!server) return fail(GRGL_ERR_MISUSE, "null argument");
    return finish(S(server), request_id, latency_ms, false);
}

GRGL_EXPORT grgl_status grgl_hip_server_wait(grgl_server* server, uint64_t request_id, double* latency_ms) {
    if (!server) return fail(GRGL_ERR_MISUSE, "null argument");
    return finish(S(server), request_id, latency_ms, true);
}

GRGL_EXPORT grgl_status grgl_hip_server_cancel(grgl_server* server, uint64_t request_id) {
    if (!server) return fail(GRGL_ERR_MISUSE, "null argument");
    Server* srv = S(server);
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->byId.find(request_id);
    if (it == srv->byId.end()) return fail(GRGL_ERR_MISUSE, "unknown request id");
    if (it->second->state != ReqState::Queued) return fail(GRGL_ERR_NOT_READY, "request already started or finished");
    it->second->state = ReqState::Cancelled;
    srv->stats.cancelled++;
    srv->queue.erase(std::remove(srv->queue.begin(), srv->queue.end(), it->second), srv->queue.end());
    srv->cvDone.notify_all();
    return GRGL_OK;
}

GRGL_EXPORT grgl_status grgl_hip_server_stats(grgl_server* server, grgl_server_stats* stats, int reset_latencies) {
    if (!server || !stats) return fail(GRGL_ERR_MISUSE, "null argument");
    Server* srv = S(server);
    std::lock_guard<std::mutex> lk(srv->mu);
    *stats = srv->stats;
    stats->hbm_in_use = srv->hbmInUse;
    stats->hbm_peak = srv->hbmPeak;
    stats->queue_depth = (uint32_t)srv->queue.size();
    stats->latency_p50_ms = percentile(srv->latencies, 0.50);
    stats->latency_p99_ms = percentile(srv->latencies, 0.99);
    double sum = 0.0;
    for (double x : srv->latencies) sum += x;
    stats->latency_mean_ms = srv->latencies.empty() ? 0.0 : sum / (double)srv->latencies.size();
    if (reset_latencies) {
        srv->latencies.clear();
        srv->hbmPeak = srv->hbmInUse;
    }
    return GRGL_OK;
}

GRGL_EXPORT grgl_status grgl_hip_server_destroy(grgl_server* server) {
    if (!server) return GRGL_OK;
    Server* srv = S(server);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        for (auto& r : srv->queue) {
            if (r->state == ReqState::Queued) {
                r->state = ReqState::Cancelled;
                srv->stats.cancelled++;
            }
        }
        srv->queue.clear();
    }
    srv->cvWork.notify_all();
    srv->cvDone.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    (void)hipSetDevice(srv->cfg.device);
    for (auto& s : srv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.raw) (void)hipFree(s.raw);
        if (s.pinned) (void)hipHostFree(s.pinned);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    delete srv;
    return GRGL_OK;
}
