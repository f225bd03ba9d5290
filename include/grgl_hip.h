/*
 * grgl_hip.h — C ABI of the MI355X (gfx950) GRG matrix multiplication library
 * (libgrgl_hip.so).
 *
 * Plain C: no HIP, torch or C++ types cross this boundary, no exceptions escape it.
 * Every entry point returns a grgl_status; on failure grgl_hip_last_error() returns a
 * thread-local message.  Streams are passed as `void*` holding a hipStream_t.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   grgl_hip_graph_create   <- convertGRGToGPUGRG (include/grgl/cuda/gpu_grg.h:747-749,
 *                              src/cuda/gpu_grg.cu:143-186): uploads an ordered GRG
 *                              (a CSRGRG, include/grgl/grg.h:1232-1300) to the device.
 *   grgl_hip_matmul         <- GRG::matrixMultiplication<IOType,NodeValueType,useBitVector>
 *                              (include/grgl/grg.h:656-667, :1342-1517), which backs
 *                              pygrgl.matmul (src/python/_grgl.cpp:100-241, :1159-1230) and
 *                              GPUGRG::matrixMultiplication (src/cuda/cuda_matmul.cu:240-408).
 *   grgl_hip_matmul_batch   <- no single reference counterpart: several matrixMultiplication
 *                              calls on one graph and direction fused into one traversal.
 *   grgl_hip_server_*       <- no reference counterpart (the reference serves concurrent
 *                              callers by releasing the GIL, src/python/_grgl.cpp:169-183).
 */
#ifndef GRGL_HIP_H
#define GRGL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/grgl/grgnode.h:57-58 */
#define GRGL_INVALID_NODE 0x7fffffffU
/* include/grgl/node_data.h:31 */
#define GRGL_COAL_COUNT_NOT_SET 0xffffffffU

typedef enum grgl_status {
    GRGL_OK = 0,
    GRGL_ERR_MISUSE = 1,      /* ApiMisuseFailure (include/grgl/common.h:45-49) */
    GRGL_ERR_DEVICE = 2,      /* HIP runtime or launch failure */
    GRGL_ERR_NOMEM = 3,       /* device allocation failed or HBM budget exceeded */
    GRGL_ERR_UNSUPPORTED = 4, /* valid request this build does not implement */
    GRGL_ERR_BUSY = 5,        /* server queue full */
    GRGL_ERR_NOT_READY = 6,   /* request still in flight */
    GRGL_ERR_CANCELLED = 7    /* request cancelled before it ran */
} grgl_status;

/* include/grgl/visitor.h:26 — same numeric values */
typedef enum grgl_direction { GRGL_DIRECTION_DOWN = 1, GRGL_DIRECTION_UP = 2 } grgl_direction;

/* The dtypes pygrgl.matmul dispatches on (src/python/_grgl.cpp:205-240). */
typedef enum grgl_dtype {
    GRGL_F64 = 0, GRGL_F32 = 1, GRGL_I64 = 2, GRGL_I32 = 3, GRGL_I16 = 4, GRGL_I8 = 5,
    GRGL_U64 = 6, GRGL_U32 = 7, GRGL_U16 = 8, GRGL_U8 = 9,
    GRGL_BOOL = 10 /* GF(2): one byte per value, XOR (include/grgl/internal_mult.h:111-128) */
} grgl_dtype;

/* GRG::NodeInitEnum (include/grgl/grg.h:646-651) — same numeric values */
typedef enum grgl_init_mode {
    GRGL_INIT_ZERO = 0,
    GRGL_INIT_VECTOR = 1, /* init[k] added to every node, row k */
    GRGL_INIT_MATRIX = 2, /* init[k * num_nodes + node] */
    GRGL_INIT_XTX = 3     /* 2 * individual coalescence count of the node */
} grgl_init_mode;

/* grgl_hip_matmul flags */
#define GRGL_FLAG_HOST_PTRS 0x1U        /* input/output/init/miss are host memory; the call is synchronous */
#define GRGL_FLAG_ACCUMULATE 0x2U       /* output (and UP miss) += result, as the reference does; else = result */
#define GRGL_FLAG_ORDER_PRESERVING 0x4U /* every node summed sequentially in the reference's order, so float
                                           results are bit-identical to the CPU path (slower on high fan-in) */

/*
 * An immutable, ordered GRG in plain arrays — the content of a CSRGRG:
 * nodes 0..num_samples-1 are the samples; every child ID is smaller than its parent's ID;
 * each node's children are sorted ascending (include/grgl/csr_storage.h:580).
 * The library copies what it needs; the caller keeps ownership of these arrays.
 */
typedef struct grgl_graph_desc {
    uint32_t num_samples;
    uint32_t num_nodes;
    uint32_t ploidy;
    uint32_t num_mutations;
    const uint64_t* down_offsets;        /* num_nodes + 1 */
    const uint32_t* down_children;       /* down_offsets[num_nodes] entries */
    const uint32_t* mutation_nodes;      /* per MutationId: node or GRGL_INVALID_NODE */
    const uint32_t* mutation_miss_nodes; /* per MutationId: missingness node or invalid; NULL if none */
    const uint32_t* coal_counts;         /* per node (samples ignored); NULL if not available */
} grgl_graph_desc;

typedef struct grgl_graph grgl_graph;

typedef struct grgl_graph_info {
    uint32_t num_samples;
    uint32_t num_nodes;
    uint32_t num_mutations;
    uint32_t ploidy;
    uint64_t num_edges;
    uint32_t up_levels;        /* topological levels of the UP (samples -> mutations) sweep */
    uint32_t down_levels;      /* topological levels of the DOWN sweep */
    uint32_t max_fanin_up;     /* largest per-node gather list, UP */
    uint32_t max_fanin_down;   /* largest per-node gather list, DOWN (parents + seeds) */
    uint64_t up_gather_entries;   /* total gather-list entries, UP */
    uint64_t down_gather_entries; /* total gather-list entries, DOWN */
    uint64_t device_bytes;     /* resident index memory on the device */
    int device;
} grgl_graph_info;

grgl_status grgl_hip_graph_create(const grgl_graph_desc* desc, int device, grgl_graph** out);
grgl_status grgl_hip_graph_destroy(grgl_graph* graph);
grgl_status grgl_hip_graph_info(const grgl_graph* graph, grgl_graph_info* info);

/*
 * Levelized-graph files: a device graph's level-bucketed plans (both directions) on disk, so
 * that a large graph is levelized once and later loaded straight to the device.  The file
 * format is this library's own (magic "GRGLLVLZ", version, sizes, arrays, 64-bit checksum);
 * loading checks the checksum and every structural invariant the kernels rely on, and
 * rejects a file that fails with GRGL_ERR_MISUSE.
 *   grgl_hip_graph_store  <- storeGPUGRGToDisk (include/grgl/cuda/gpu_grg.h:473-480,
 *                            src/cuda/gpu_grg.cu:10-68): writes a device graph's plans.
 *   grgl_hip_graph_load   <- loadGPUGRGFromDisk (include/grgl/cuda/gpu_grg.h:482-489,
 *                            src/cuda/gpu_grg.cu:70-141): a device graph from such a file,
 *                            as grgl_hip_graph_create would have built it.
 *   grgl_hip_levelize_to_file <- no reference counterpart (the reference levelizes only on
 *                            a machine with a GPU): levelizes a graph on the host, no device
 *                            needed, and writes the file.
 *   grgl_hip_levelized_file_info <- no reference counterpart: validates a file on the host
 *                            and reports its sizes (device = -1, device_bytes = what a load
 *                            would allocate).
 */
grgl_status grgl_hip_graph_store(const grgl_graph* graph, const char* path);
grgl_status grgl_hip_graph_load(const char* path, int device, grgl_graph** out);
grgl_status grgl_hip_levelize_to_file(const grgl_graph_desc* desc, const char* path);
grgl_status grgl_hip_levelized_file_info(const char* path, grgl_graph_info* info);

/*
 * GRG::matrixMultiplication on the device, argument for argument
 * (include/grgl/grg.h:656-667), plus dtype, flags and a stream:
 *   input        input_rows (K) x input_cols, row-major.  UP: input_cols = N (or N/ploidy
 *                with by_individual).  DOWN: input_cols = M.
 *   output       output_size = K x output_cols, row-major.  UP: M columns, DOWN: N (or
 *                N/ploidy), emit_all_nodes: num_nodes.
 *   init         K values (VECTOR) or K x num_nodes (MATRIX); NULL otherwise.
 *   miss         K x M.  DOWN: an input added into each mutation's missingness node.
 *                UP: an output, accumulated like `output`.
 * Without GRGL_FLAG_HOST_PTRS every pointer is device memory and the work is enqueued on
 * `stream` (a hipStream_t; NULL = the device's null stream) without a host sync.
 * Any K: inputs of more than 65536 rows, or whose row buffer would take more than half of
 * the free device memory, run as several traversals of consecutive row blocks (rows are
 * independent; results are the same).
 */
grgl_status grgl_hip_matmul(grgl_graph* graph, grgl_dtype dtype, const void* input, size_t input_cols,
                            size_t input_rows, grgl_direction direction, void* output, size_t output_size,
                            int emit_all_nodes, int by_individual, const void* init, grgl_init_mode init_mode,
                            void* miss, uint32_t flags, void* stream);

/*
 * grgl_hip_matmul without any device allocation, for callers that cannot allocate: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  No counterpart in
 * the reference.  Same arguments, checks and results (bit for bit) as grgl_hip_matmul with
 * device pointers; GRGL_FLAG_HOST_PTRS is rejected (GRGL_ERR_MISUSE).  The traversal's row
 * buffer lives in a device array of GRGL_NOALLOC_SCRATCH_BYTES that the library holds on every
 * GPU it runs on until the process ends; inputs of more rows than fit are run as several
 * traversals of row blocks, and a graph whose single-row buffer exceeds the array is rejected
 * with GRGL_ERR_MISUSE.  The row buffer is (num_nodes [+ num_mutations, or 2 x num_mutations
 * with missing data, DOWN]) rows of K x dtype size bytes rounded up to 4, 8, 16 bytes and then
 * to a power of two under 64 bytes, to 64 bytes up to 256, to 256 bytes above.
 * Every call on one GPU shares that array: calls that overlap on one GPU (from two threads, or
 * on two streams with no order between them) race and corrupt each other's results.  The call
 * enqueues kernels only (no hipGraph, no events; kernel timing does not see it) and returns
 * without a host sync.
 */
#define GRGL_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
grgl_status grgl_hip_matmul_noalloc(grgl_graph* graph, grgl_dtype dtype, const void* input, size_t input_cols,
                                    size_t input_rows, grgl_direction direction, void* output, size_t output_size,
                                    int emit_all_nodes, int by_individual, const void* init, grgl_init_mode init_mode,
                                    void* miss, uint32_t flags, void* stream);

/*
 * Several requests on one graph, one direction and one dtype carried by ONE traversal:
 * their rows are stacked (each request starting on a 16-byte boundary of the row; stacked
 * rows <= 65536) and share every edge read.  Each request keeps its own input/output
 * buffers and gets bit-for-bit the result of a grgl_hip_matmul call with the same flags:
 * every column is summed by the same arithmetic whatever else is stacked beside it.
 * No init/miss/emit_all/by_individual here.
 */
typedef struct grgl_request {
    uint32_t rows;
    const void* input;  /* rows x input_cols */
    void* output;       /* rows x output_cols */
} grgl_request;

grgl_status grgl_hip_matmul_batch(grgl_graph* graph, grgl_dtype dtype, grgl_direction direction,
                                  const grgl_request* requests, uint32_t num_requests, uint32_t flags,
                                  void* stream);

/* Tuning of the level kernels for (graph, direction, dtype, rows).  0 = library default.
 * Team width, hipGraph replay and block order only change speed.  heavy_fanin changes which
 * gather lists are split 64 ways (float rounding of those rows) for every later call in that
 * direction.  heavy_fanin and max_launch_entries (and grgl_hip_tuning_load) replace the
 * direction's launch table; a traversal already enqueued or being enqueued on another thread
 * keeps the table it started with, so retuning is safe while traversals run (only calls that
 * start afterwards see the new heavy threshold). */
typedef struct grgl_tuning {
    uint32_t team_lanes;     /* lanes that share one node's row slice: a power of two up to the
                                row's 16-byte slices rounded up, or up to 8 for short rows (which
                                then take 8- or 4-byte accesses and idle lanes); larger values are
                                clamped */
    uint32_t heavy_fanin;    /* gather lists longer than this are split into 64 partials */
    uint32_t use_graph;      /* 1 = replay the level sequence as a hipGraph (built once per row buffer and
                                shape from explicit kernel nodes, cached per graph) */
    uint32_t xcd_map;        /* block order of the level kernels when the row splits into a multiple
                                of 8 column slices: 0 = default (plain 2-D grid), 1 = XCD-aware (XCD i
                                takes slices i, i+8, ...), 2 = plain 2-D grid.  The plain grid runs the
                                column slices one after another, so one slice of every row (V x 256 B)
                                stays in the memory-side cache shared by all XCDs; measured faster. */
    uint32_t max_launch_entries; /* gather entries per level-kernel launch at most (0 = library default,
                                    2^30, also the upper limit); larger levels are split into several
                                    launches.  Never changes results; like heavy_fanin it applies to
                                    every later call in that direction. */
    uint32_t chain;          /* form of the level sequence; never changes results.
                                0 = default: grouped levels (3) where the graph has them for rows of
                                at most 256 bytes (DOWN) or 512 bytes (UP), else 1.
                                1 = one launch per level.
                                2 = the whole level sequence as ONE chained launch where the shape
                                allows it (rows of at most 64 bytes of a 4- or 8-byte type, at most
                                120 levels): blocks take the levels' work in order from a ticket
                                counter and wait on per-level completion counters, with no grid
                                barrier.  Its waits are bounded (grgl_hip_chain_control).
                                3 = grouped levels: not order-preserving, with the default
                                heavy_fanin.  The later levels of the direction are cut, when the
                                graph is created, into connected groups of rows that read no row of
                                those levels outside the group; ONE launch runs them, each block its
                                groups level after level from LDS, with no wait on another block
                                (rows wider than 64 bytes as 64-byte column slices, one grid row
                                each).  The levels before run one launch each. */
    uint32_t block_threads;  /* threads per block: of the level kernels with wide teams (16-byte
                                accesses, 8-32 lanes per row) of 4- and 8-byte types, 256 (default) or
                                512; of grouped launches with rows of at most 8 bytes, 1024 (default)
                                or 256 (wider rows: 256).  A value a launch does not take leaves it at
                                its default.  Never changes results. */
    uint32_t loads_in_flight; /* row loads issued before the first add in each batch of such kernels:
                                0 = default (4), 4 or 8.  Never changes results. */
} grgl_tuning;

grgl_status grgl_hip_set_tuning(grgl_graph* graph, grgl_direction direction, grgl_dtype dtype, uint32_t rows,
                                const grgl_tuning* tuning);
/* Chained traversals (grgl_tuning::chain = 2) on `graph`: every wait of one is bounded by
 * wait_us (default 200000 us; 0 leaves it unchanged) and gives up when the abort word is set
 * (abort = 1 sets it, 0 clears it, -1 leaves it; blocks read it when they start and while they
 * wait).  A traversal that gave up stops and leaves its outputs undefined, and the graph is
 * marked: *gave_up (if not NULL) receives 1 when marked, 0 otherwise; reset = 1 clears the mark.
 * While marked, every matmul call on the graph fails with GRGL_ERR_DEVICE (a call with host
 * buffers also checks after its own traversal).  No counterpart in the reference. */
grgl_status grgl_hip_chain_control(grgl_graph* graph, int abort, uint32_t wait_us, uint32_t* gave_up, int reset);
/* Time candidate tunings on the device and keep the fastest; writes it to *chosen and its time
 * to *best_ms: of a whole traversal (seed, levels, read-out) of zero inputs for rows of at
 * most 512 bytes (rows * dtype size) whose input and output fit 4 GiB, else of the level
 * sequence alone.  While it runs, other calls
 * of the same (direction, dtype, rows) may take any candidate (results never depend on it). */
grgl_status grgl_hip_autotune(grgl_graph* graph, grgl_direction direction, grgl_dtype dtype, uint32_t rows,
                              grgl_tuning* chosen, float* best_ms);
/* The tuning a call of (direction, dtype, rows) uses now: the stored choice with library
 * defaults filled in (team_lanes as launched, use_graph, xcd_map; chain as taken: 3 only where
 * the direction has grouped levels for this row size, else 1; block_threads of the grouped
 * launch when chain is 3, else of the level kernels) and the direction's heavy_fanin and
 * max_launch_entries. */
grgl_status grgl_hip_get_tuning(grgl_graph* graph, grgl_direction direction, grgl_dtype dtype, uint32_t rows,
                                grgl_tuning* out);
/* Tuning cache file (text).  save writes every stored (direction, dtype class, rows) choice
 * and each direction's heavy_fanin / max_launch_entries with a fingerprint of the graph's
 * level structure; load applies a file written for a graph of the same structure and fails
 * with GRGL_ERR_MISUSE for any other graph or a malformed file.  Autotuning once and loading
 * the file in later processes skips the timing runs. */
grgl_status grgl_hip_tuning_save(grgl_graph* graph, const char* path);
grgl_status grgl_hip_tuning_load(grgl_graph* graph, const char* path);

/*
 * Per-launch timing of the level (pull-gather) kernels.  When enabled, each level launch
 * of subsequent calls is bracketed by HIP events on the stream it runs on; the call below
 * syncs and reports the summed kernel milliseconds, launch count and the algorithmic
 * bytes those launches moved (4 B per gather-list entry and per node offset, plus one
 * row read per entry and one row write per node).
 */
grgl_status grgl_hip_kernel_timing(grgl_graph* graph, int enable);
grgl_status grgl_hip_kernel_timing_read(grgl_graph* graph, double* total_ms, uint64_t* launches,
                                        double* algorithmic_bytes, int reset);

/* ----------------------------------------------------------------------------------------
 * Serving: a resident-graph request loop with continuous batching.  Requests on the same
 * (graph, direction, dtype) that are queued when a stream slot frees up are stacked into
 * one traversal; admission is bounded by an HBM budget for in-flight work.
 * Host-memory requests; results are written to the caller's output buffer before the
 * request completes.
 * -------------------------------------------------------------------------------------- */
typedef struct grgl_server grgl_server;

typedef struct grgl_server_config {
    int device;
    uint32_t num_streams;      /* concurrent traversals on the device (slots); default 2 */
    uint32_t max_batch_rows;   /* stacked rows per traversal; default 512 */
    uint64_t hbm_budget_bytes; /* device bytes for in-flight batches; default 8 GiB */
    uint32_t max_queue;        /* queued requests before submit returns BUSY; default 4096 */
    uint32_t flags;            /* GRGL_FLAG_ORDER_PRESERVING allowed */
    uint32_t batch_window_us;  /* a free slot waits up to this long after the oldest queued request
                                  arrived, to stack more requests into its batch; default 0 */
} grgl_server_config;

typedef struct grgl_server_stats {
    uint64_t submitted;
    uint64_t completed;
    uint64_t rejected;
    uint64_t cancelled;
    uint64_t batches;
    uint64_t rows_processed;
    uint64_t edge_rhs;          /* sum over batches of gather entries x rows */
    uint64_t hbm_in_use;        /* bytes held by in-flight batches now */
    uint64_t hbm_peak;
    uint32_t queue_depth;
    double latency_p50_ms;      /* submit -> completion, over completed requests */
    double latency_p99_ms;
    double latency_mean_ms;
} grgl_server_stats;

grgl_status grgl_hip_server_create(const grgl_server_config* config, grgl_server** out);
grgl_status grgl_hip_server_add_graph(grgl_server* server, grgl_graph* graph, uint32_t* graph_id);
grgl_status grgl_hip_server_submit(grgl_server* server, uint32_t graph_id, grgl_dtype dtype,
                                   grgl_direction direction, uint32_t rows, const void* host_input,
                                   void* host_output, uint64_t* request_id);
/* Same, with device buffers (any device allocation on the server's device): no staging
 * copies; the traversal reads the input and writes the output in place.  Both buffers
 * must stay valid until the request completes. */
grgl_status grgl_hip_server_submit_device(grgl_server* server, uint32_t graph_id, grgl_dtype dtype,
                                          grgl_direction direction, uint32_t rows, const void* device_input,
                                          void* device_output, uint64_t* request_id);
/* GRGL_OK when done (result written), GRGL_ERR_NOT_READY while queued or running. */
grgl_status grgl_hip_server_poll(grgl_server* server, uint64_t request_id, double* latency_ms);
grgl_status grgl_hip_server_wait(grgl_server* server, uint64_t request_id, double* latency_ms);
/* Cancels a request that has not started; GRGL_ERR_NOT_READY if it is already running. */
grgl_status grgl_hip_server_cancel(grgl_server* server, uint64_t request_id);
grgl_status grgl_hip_server_stats(grgl_server* server, grgl_server_stats* stats, int reset_latencies);
grgl_status grgl_hip_server_destroy(grgl_server* server);

/* Thread-local message for the last non-OK status returned on this thread. */
const char* grgl_hip_last_error(void);
/* Number of visible devices (0 when no GPU is present). */
int grgl_hip_device_count(void);

/* Page-locked host memory (hipHostMalloc) for callers that stage GRGL_FLAG_HOST_PTRS buffers:
 * copies to and from it run at full link speed.  Free with grgl_hip_host_free.  No
 * counterpart in the reference. */
grgl_status grgl_hip_host_alloc(size_t bytes, void** out);
grgl_status grgl_hip_host_free(void* ptr);

/* Measured streaming bandwidth of `device`, in GB/s: a kernel reading `bytes` (at least 4096,
 * under 32 GiB) with one 16-byte load per thread, and a kernel copying them (read + write
 * bytes counted).  Allocates and frees two buffers of `bytes`.  No counterpart in the
 * reference: it gives the practical ceiling the level kernels' HBM traffic is compared with. */
grgl_status grgl_hip_stream_ceiling(int device, size_t bytes, double* read_gbps, double* copy_gbps);
/* Measured random-gather rate of `device`, in 4-byte gathers per second: rows of 16 list
 * entries, each entry an index (read in order) of a value gathered from a table of
 * `table_bytes` at a random position, one working lane in `lanes_per_row` (1 or 4).  The
 * access of a traversal at one column (4-byte rows), whose levels are bound by it.  No
 * counterpart in the reference. */
grgl_status grgl_hip_gather_ceiling(int device, size_t table_bytes, uint32_t lanes_per_row, double* gathers_per_s);

#ifdef __cplusplus
}
#endif
#endif /* GRGL_HIP_H */
