"""ctypes view of the C ABI in include/grgl_hip.h (libgrgl_hip.so).

This is the boundary a non-Python host (or a test) binds: plain pointers and sizes.
Device pointers may come from any allocator (torch tensors' data_ptr() are used by the
tests and the benchmark); host pointers are accepted with FLAG_HOST_PTRS.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrgl_hip.so")

OK = 0
ERR_MISUSE = 1
ERR_DEVICE = 2
ERR_NOMEM = 3
ERR_UNSUPPORTED = 4
ERR_BUSY = 5
ERR_NOT_READY = 6
ERR_CANCELLED = 7

DOWN = 1
UP = 2

F64, F32, I64, I32, I16, I8, U64, U32, U16, U8, BOOL = range(11)
INIT_ZERO, INIT_VECTOR, INIT_MATRIX, INIT_XTX = 0, 1, 2, 3

FLAG_HOST_PTRS = 0x1
FLAG_ACCUMULATE = 0x2
FLAG_ORDER_PRESERVING = 0x4
# Size of the device array holding the row buffer of grgl_hip_matmul_noalloc (GRGL_NOALLOC_SCRATCH_BYTES).
NOALLOC_SCRATCH_BYTES = 128 << 20

INVALID_NODE = 0x7FFFFFFF
COAL_COUNT_NOT_SET = 0xFFFFFFFF

NUMPY_TO_DTYPE = {
    np.dtype(np.float64): F64,
    np.dtype(np.float32): F32,
    np.dtype(np.int64): I64,
    np.dtype(np.int32): I32,
    np.dtype(np.int16): I16,
    np.dtype(np.int8): I8,
    np.dtype(np.uint64): U64,
    np.dtype(np.uint32): U32,
    np.dtype(np.uint16): U16,
    np.dtype(np.uint8): U8,
    np.dtype(np.bool_): BOOL,
}

# Every entry point include/grgl_hip.h declares.
EXPORTED_SYMBOLS = [
    "grgl_hip_graph_create",
    "grgl_hip_graph_destroy",
    "grgl_hip_graph_info",
    "grgl_hip_graph_store",
    "grgl_hip_graph_load",
    "grgl_hip_levelize_to_file",
    "grgl_hip_levelized_file_info",
    "grgl_hip_matmul",
    "grgl_hip_matmul_noalloc",
    "grgl_hip_matmul_batch",
    "grgl_hip_set_tuning",
    "grgl_hip_autotune",
    "grgl_hip_get_tuning",
    "grgl_hip_chain_control",
    "grgl_hip_tuning_save",
    "grgl_hip_tuning_load",
    "grgl_hip_kernel_timing",
    "grgl_hip_kernel_timing_read",
    "grgl_hip_server_create",
    "grgl_hip_server_add_graph",
    "grgl_hip_server_submit",
    "grgl_hip_server_submit_device",
    "grgl_hip_server_poll",
    "grgl_hip_server_wait",
    "grgl_hip_server_cancel",
    "grgl_hip_server_stats",
    "grgl_hip_server_destroy",
    "grgl_hip_last_error",
    "grgl_hip_device_count",
    "grgl_hip_host_alloc",
    "grgl_hip_host_free",
    "grgl_hip_stream_ceiling",
    "grgl_hip_gather_ceiling",
]


class GraphDesc(ctypes.Structure):
    _fields_ = [
        ("num_samples", ctypes.c_uint32),
        ("num_nodes", ctypes.c_uint32),
        ("ploidy", ctypes.c_uint32),
        ("num_mutations", ctypes.c_uint32),
        ("down_offsets", ctypes.c_void_p),
        ("down_children", ctypes.c_void_p),
        ("mutation_nodes", ctypes.c_void_p),
        ("mutation_miss_nodes", ctypes.c_void_p),
        ("coal_counts", ctypes.c_void_p),
    ]


class GraphInfo(ctypes.Structure):
    _fields_ = [
        ("num_samples", ctypes.c_uint32),
        ("num_nodes", ctypes.c_uint32),
        ("num_mutations", ctypes.c_uint32),
        ("ploidy", ctypes.c_uint32),
        ("num_edges", ctypes.c_uint64),
        ("up_levels", ctypes.c_uint32),
        ("down_levels", ctypes.c_uint32),
        ("max_fanin_up", ctypes.c_uint32),
        ("max_fanin_down", ctypes.c_uint32),
        ("up_gather_entries", ctypes.c_uint64),
        ("down_gather_entries", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
        ("device", ctypes.c_int),
    ]


class Request(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_uint32), ("input", ctypes.c_void_p), ("output", ctypes.c_void_p)]


class Tuning(ctypes.Structure):
    _fields_ = [("team_lanes", ctypes.c_uint32), ("heavy_fanin", ctypes.c_uint32), ("use_graph", ctypes.c_uint32),
                ("xcd_map", ctypes.c_uint32), ("max_launch_entries", ctypes.c_uint32), ("chain", ctypes.c_uint32),
                ("block_threads", ctypes.c_uint32), ("loads_in_flight", ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int),
        ("num_streams", ctypes.c_uint32),
        ("max_batch_rows", ctypes.c_uint32),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("max_queue", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("batch_window_us", ctypes.c_uint32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
        ("cancelled", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
        ("rows_processed", ctypes.c_uint64),
        ("edge_rhs", ctypes.c_uint64),
        ("hbm_in_use", ctypes.c_uint64),
        ("hbm_peak", ctypes.c_uint64),
        ("queue_depth", ctypes.c_uint32),
        ("latency_p50_ms", ctypes.c_double),
        ("latency_p99_ms", ctypes.c_double),
        ("latency_mean_ms", ctypes.c_double),
    ]


_lib = None


def lib():
    """Load libgrgl_hip.so (raises OSError if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    from ._runtime import share_hip_runtime_with_torch

    share_hip_runtime_with_torch()
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, sz, ci = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int
    L.grgl_hip_last_error.restype = ctypes.c_char_p
    L.grgl_hip_last_error.argtypes = []
    L.grgl_hip_device_count.restype = ci
    L.grgl_hip_device_count.argtypes = []
    L.grgl_hip_host_alloc.restype = ci
    L.grgl_hip_host_alloc.argtypes = [sz, ctypes.POINTER(vp)]
    L.grgl_hip_host_free.restype = ci
    L.grgl_hip_host_free.argtypes = [vp]
    L.grgl_hip_stream_ceiling.restype = ci
    L.grgl_hip_stream_ceiling.argtypes = [ci, sz, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.grgl_hip_gather_ceiling.restype = ci
    L.grgl_hip_gather_ceiling.argtypes = [ci, sz, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    L.grgl_hip_graph_create.restype = ci
    L.grgl_hip_graph_create.argtypes = [ctypes.POINTER(GraphDesc), ci, ctypes.POINTER(vp)]
    L.grgl_hip_graph_destroy.restype = ci
    L.grgl_hip_graph_destroy.argtypes = [vp]
    L.grgl_hip_graph_info.restype = ci
    L.grgl_hip_graph_info.argtypes = [vp, ctypes.POINTER(GraphInfo)]
    L.grgl_hip_graph_store.restype = ci
    L.grgl_hip_graph_store.argtypes = [vp, ctypes.c_char_p]
    L.grgl_hip_graph_load.restype = ci
    L.grgl_hip_graph_load.argtypes = [ctypes.c_char_p, ci, ctypes.POINTER(vp)]
    L.grgl_hip_levelize_to_file.restype = ci
    L.grgl_hip_levelize_to_file.argtypes = [ctypes.POINTER(GraphDesc), ctypes.c_char_p]
    L.grgl_hip_levelized_file_info.restype = ci
    L.grgl_hip_levelized_file_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(GraphInfo)]
    L.grgl_hip_matmul.restype = ci
    L.grgl_hip_matmul.argtypes = [vp, ci, vp, sz, sz, ci, vp, sz, ci, ci, vp, ci, vp, u32, vp]
    if hasattr(L, "grgl_hip_matmul_noalloc"):
        L.grgl_hip_matmul_noalloc.restype = ci
        L.grgl_hip_matmul_noalloc.argtypes = [vp, ci, vp, sz, sz, ci, vp, sz, ci, ci, vp, ci, vp, u32, vp]
    L.grgl_hip_matmul_batch.restype = ci
    L.grgl_hip_matmul_batch.argtypes = [vp, ci, ci, ctypes.POINTER(Request), u32, u32, vp]
    L.grgl_hip_set_tuning.restype = ci
    L.grgl_hip_set_tuning.argtypes = [vp, ci, ci, u32, ctypes.POINTER(Tuning)]
    L.grgl_hip_kernel_timing.restype = ci
    L.grgl_hip_kernel_timing.argtypes = [vp, ci]
    L.grgl_hip_kernel_timing_read.restype = ci
    L.grgl_hip_kernel_timing_read.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64),
                                              ctypes.POINTER(ctypes.c_double), ci]
    if hasattr(L, "grgl_hip_autotune"):
        L.grgl_hip_autotune.restype = ci
        L.grgl_hip_autotune.argtypes = [vp, ci, ci, u32, ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_float)]
        L.grgl_hip_chain_control.restype = ci
        L.grgl_hip_chain_control.argtypes = [vp, ci, u32, ctypes.POINTER(u32), ci]
        L.grgl_hip_get_tuning.restype = ci
        L.grgl_hip_get_tuning.argtypes = [vp, ci, ci, u32, ctypes.POINTER(Tuning)]
        L.grgl_hip_tuning_save.restype = ci
        L.grgl_hip_tuning_save.argtypes = [vp, ctypes.c_char_p]
        L.grgl_hip_tuning_load.restype = ci
        L.grgl_hip_tuning_load.argtypes = [vp, ctypes.c_char_p]
    if hasattr(L, "grgl_hip_server_create"):
        L.grgl_hip_server_create.restype = ci
        L.grgl_hip_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
        L.grgl_hip_server_add_graph.restype = ci
        L.grgl_hip_server_add_graph.argtypes = [vp, vp, ctypes.POINTER(u32)]
        L.grgl_hip_server_submit.restype = ci
        L.grgl_hip_server_submit.argtypes = [vp, u32, ci, ci, u32, vp, vp, ctypes.POINTER(u64)]
        L.grgl_hip_server_submit_device.restype = ci
        L.grgl_hip_server_submit_device.argtypes = [vp, u32, ci, ci, u32, vp, vp, ctypes.POINTER(u64)]
        L.grgl_hip_server_poll.restype = ci
        L.grgl_hip_server_poll.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_double)]
        L.grgl_hip_server_wait.restype = ci
        L.grgl_hip_server_wait.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_double)]
        L.grgl_hip_server_cancel.restype = ci
        L.grgl_hip_server_cancel.argtypes = [vp, u64]
        L.grgl_hip_server_stats.restype = ci
        L.grgl_hip_server_stats.argtypes = [vp, ctypes.POINTER(ServerStats), ci]
        L.grgl_hip_server_destroy.restype = ci
        L.grgl_hip_server_destroy.argtypes = [vp]
    _lib = L
    return L


class GrglError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"grgl_hip status {status}: {message}")
        self.status = status


def check(status):
    if status != OK:
        raise GrglError(status, lib().grgl_hip_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


def stream_ceiling(device=0, nbytes=2 << 30):
    """grgl_hip_stream_ceiling: measured streaming read and copy bandwidth of the device in
    GB/s, as {"read_GBps": r, "copy_GBps": c}."""
    r, c = ctypes.c_double(), ctypes.c_double()
    check(lib().grgl_hip_stream_ceiling(device, nbytes, ctypes.byref(r), ctypes.byref(c)))
    return {"read_GBps": r.value, "copy_GBps": c.value}


def gather_ceiling(device=0, table_bytes=2 << 20, lanes_per_row=1):
    """grgl_hip_gather_ceiling: measured random 4-byte gathers per second of the device."""
    g = ctypes.c_double()
    check(lib().grgl_hip_gather_ceiling(device, table_bytes, lanes_per_row, ctypes.byref(g)))
    return g.value


def _desc(arrays):
    """A GraphDesc over contiguous copies of the arrays (returned with the copies to keep)."""
    keep = {
        "off": np.ascontiguousarray(arrays["down_offsets"], dtype=np.uint64),
        "child": np.ascontiguousarray(arrays["down_children"], dtype=np.uint32),
        "mut": np.ascontiguousarray(arrays["mutation_nodes"], dtype=np.uint32),
    }
    miss = arrays.get("mutation_miss_nodes")
    coal = arrays.get("coal_counts")
    if miss is not None:
        keep["miss"] = np.ascontiguousarray(miss, dtype=np.uint32)
    if coal is not None:
        keep["coal"] = np.ascontiguousarray(coal, dtype=np.uint32)
    d = GraphDesc()
    d.num_samples = int(arrays["num_samples"])
    d.num_nodes = len(keep["off"]) - 1
    d.ploidy = int(arrays["ploidy"])
    d.num_mutations = len(keep["mut"])
    d.down_offsets = _ptr(keep["off"])
    d.down_children = _ptr(keep["child"])
    d.mutation_nodes = _ptr(keep["mut"])
    d.mutation_miss_nodes = _ptr(keep.get("miss"))
    d.coal_counts = _ptr(keep.get("coal"))
    return d, keep


def levelize_to_file(arrays, path):
    """Levelize a graph on the host (no device needed) into a levelized-graph file that
    Graph.load reads (grgl_hip_levelize_to_file)."""
    d, keep = _desc(arrays)
    check(lib().grgl_hip_levelize_to_file(ctypes.byref(d), os.fsencode(path)))


def levelized_file_info(path):
    """Validate a levelized-graph file on the host; its sizes (grgl_hip_levelized_file_info)."""
    i = GraphInfo()
    check(lib().grgl_hip_levelized_file_info(os.fsencode(path), ctypes.byref(i)))
    return {f: getattr(i, f) for f, _ in GraphInfo._fields_}


class Graph:
    """A device-resident graph created through grgl_hip_graph_create (or loaded from a
    levelized-graph file with Graph.load)."""

    @classmethod
    def load(cls, path, device=0):
        """grgl_hip_graph_load: the device graph a levelized-graph file holds."""
        h = ctypes.c_void_p()
        check(lib().grgl_hip_graph_load(os.fsencode(path), device, ctypes.byref(h)))
        self = cls.__new__(cls)
        self._keep = {}
        self.handle = h
        info = self.info()
        self.num_samples = info["num_samples"]
        self.num_nodes = info["num_nodes"]
        self.num_mutations = info["num_mutations"]
        self.ploidy = info["ploidy"]
        return self

    def store(self, path):
        """grgl_hip_graph_store: write this graph's plans to a levelized-graph file."""
        check(lib().grgl_hip_graph_store(self.handle, os.fsencode(path)))

    def __init__(self, arrays, device=0):
        d, self._keep = _desc(arrays)
        h = ctypes.c_void_p()
        check(lib().grgl_hip_graph_create(ctypes.byref(d), device, ctypes.byref(h)))
        self.handle = h
        self.num_samples = d.num_samples
        self.num_nodes = d.num_nodes
        self.num_mutations = d.num_mutations
        self.ploidy = d.ploidy

    def close(self):
        if self.handle:
            check(lib().grgl_hip_graph_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = GraphInfo()
        check(lib().grgl_hip_graph_info(self.handle, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in GraphInfo._fields_}

    def matmul_ptr(self, dtype, in_ptr, in_cols, rows, direction, out_ptr, out_size, emit_all=False,
                   by_individual=False, init_ptr=None, init_mode=INIT_ZERO, miss_ptr=None, flags=0, stream=None):
        check(lib().grgl_hip_matmul(self.handle, dtype, in_ptr, in_cols, rows, direction, out_ptr, out_size,
                                    int(emit_all), int(by_individual), init_ptr, init_mode, miss_ptr, flags, stream))

    def matmul_noalloc_ptr(self, dtype, in_ptr, in_cols, rows, direction, out_ptr, out_size, emit_all=False,
                           by_individual=False, init_ptr=None, init_mode=INIT_ZERO, miss_ptr=None, flags=0,
                           stream=None):
        """grgl_hip_matmul_noalloc: device pointers, no device allocation (row buffer in the
        library's fixed scratch array; overlapping calls on one GPU race)."""
        check(lib().grgl_hip_matmul_noalloc(self.handle, dtype, in_ptr, in_cols, rows, direction, out_ptr, out_size,
                                            int(emit_all), int(by_individual), init_ptr, init_mode, miss_ptr, flags,
                                            stream))

    def matmul_host(self, inp, direction, emit_all=False, by_individual=False, init=None, miss=None, flags=0):
        """Host numpy in, numpy out (zero-initialised, accumulated: pygrgl semantics)."""
        inp = np.ascontiguousarray(inp)
        dt = NUMPY_TO_DTYPE[inp.dtype]
        rows, cols = inp.shape
        if emit_all:
            out_cols = self.num_nodes
        elif direction == DOWN:
            out_cols = self.num_samples // self.ploidy if by_individual else self.num_samples
        else:
            out_cols = self.num_mutations
        out = np.zeros((rows, out_cols), dtype=inp.dtype)
        mode, init_arr = INIT_ZERO, None
        if isinstance(init, str):
            mode = INIT_XTX
        elif init is not None:
            init_arr = np.ascontiguousarray(init, dtype=inp.dtype)
            mode = INIT_VECTOR if init_arr.ndim == 1 else INIT_MATRIX
        self.matmul_ptr(dt, inp.ctypes.data, cols, rows, direction, out.ctypes.data, out.size, emit_all,
                        by_individual, _ptr(init_arr), mode, _ptr(miss), flags | FLAG_HOST_PTRS | FLAG_ACCUMULATE)
        return out

    def matmul_batch_ptr(self, dtype, direction, reqs, flags=0, stream=None):
        """reqs: list of (rows, in_ptr, out_ptr)."""
        arr = (Request * len(reqs))()
        for i, (r, ip, op) in enumerate(reqs):
            arr[i].rows = r
            arr[i].input = ip
            arr[i].output = op
        check(lib().grgl_hip_matmul_batch(self.handle, dtype, direction, arr, len(reqs), flags, stream))

    def set_tuning(self, direction, dtype, rows, team_lanes=0, heavy_fanin=0, use_graph=0, xcd_map=0,
                   max_launch_entries=0, chain=0, block_threads=0, loads_in_flight=0):
        t = Tuning(team_lanes, heavy_fanin, use_graph, xcd_map, max_launch_entries, chain, block_threads,
                   loads_in_flight)
        check(lib().grgl_hip_set_tuning(self.handle, direction, dtype, rows, ctypes.byref(t)))

    def autotune(self, direction, dtype, rows):
        t = Tuning()
        ms = ctypes.c_float()
        check(lib().grgl_hip_autotune(self.handle, direction, dtype, rows, ctypes.byref(t), ctypes.byref(ms)))
        return {"team_lanes": t.team_lanes, "heavy_fanin": t.heavy_fanin, "use_graph": t.use_graph,
                "xcd_map": t.xcd_map, "max_launch_entries": t.max_launch_entries, "chain": t.chain,
                "block_threads": t.block_threads, "loads_in_flight": t.loads_in_flight}, ms.value

    def get_tuning(self, direction, dtype, rows):
        t = Tuning()
        check(lib().grgl_hip_get_tuning(self.handle, direction, dtype, rows, ctypes.byref(t)))
        return {"team_lanes": t.team_lanes, "heavy_fanin": t.heavy_fanin, "use_graph": t.use_graph,
                "xcd_map": t.xcd_map, "max_launch_entries": t.max_launch_entries, "chain": t.chain,
                "block_threads": t.block_threads, "loads_in_flight": t.loads_in_flight}

    def chain_control(self, abort=-1, wait_us=0, reset=False):
        """grgl_hip_chain_control: set (1) / clear (0) / keep (-1) the abort word, bound chained
        waits (wait_us > 0), and return 1 if a chained traversal on the graph gave up (then reset)."""
        n = ctypes.c_uint32()
        check(lib().grgl_hip_chain_control(self.handle, abort, wait_us, ctypes.byref(n), int(reset)))
        return n.value

    def save_tuning(self, path):
        check(lib().grgl_hip_tuning_save(self.handle, os.fsencode(path)))

    def load_tuning(self, path):
        check(lib().grgl_hip_tuning_load(self.handle, os.fsencode(path)))

    def kernel_timing(self, enable):
        check(lib().grgl_hip_kernel_timing(self.handle, int(enable)))

    def kernel_timing_read(self, reset=True):
        ms, n, b = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_double()
        check(lib().grgl_hip_kernel_timing_read(self.handle, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b),
                                                int(reset)))
        return ms.value, n.value, b.value


class Server:
    """grgl_hip_server_*: resident graphs, continuous batching, HBM-budget admission.

    Requests are host numpy arrays; each submit returns (request_id, output array) and the
    output is filled when wait()/poll() reports completion.  Inputs and outputs are kept
    alive by this object until the request finishes.
    """

    def __init__(self, device=0, num_streams=2, max_batch_rows=512, hbm_budget_bytes=0, max_queue=0, flags=0,
                 batch_window_us=0):
        cfg = ServerConfig(device, num_streams, max_batch_rows, hbm_budget_bytes, max_queue, flags, batch_window_us)
        h = ctypes.c_void_p()
        check(lib().grgl_hip_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._graphs = []
        self._live = {}

    def add_graph(self, graph):
        gid = ctypes.c_uint32()
        check(lib().grgl_hip_server_add_graph(self.handle, graph.handle, ctypes.byref(gid)))
        self._graphs.append(graph)
        return gid.value

    def submit(self, graph_id, x, direction):
        x = np.ascontiguousarray(x)
        g = self._graphs[graph_id]
        rows = x.shape[0]
        out_cols = g.num_mutations if direction == UP else g.num_samples
        out = np.zeros((rows, out_cols), dtype=x.dtype)
        rid = ctypes.c_uint64()
        check(lib().grgl_hip_server_submit(self.handle, graph_id, NUMPY_TO_DTYPE[x.dtype], direction, rows,
                                           x.ctypes.data, out.ctypes.data, ctypes.byref(rid)))
        self._live[rid.value] = (x, out)
        return rid.value, out

    def submit_device(self, graph_id, dtype, direction, rows, in_ptr, out_ptr):
        """Device buffers (e.g. torch tensors' data_ptr()); they must outlive the request."""
        rid = ctypes.c_uint64()
        check(lib().grgl_hip_server_submit_device(self.handle, graph_id, dtype, direction, rows, in_ptr, out_ptr,
                                                  ctypes.byref(rid)))
        return rid.value

    def _finish(self, rid, fn):
        lat = ctypes.c_double()
        st = fn(self.handle, rid, ctypes.byref(lat))
        if st == ERR_NOT_READY:
            return None
        self._live.pop(rid, None)
        check(st)
        return lat.value

    def wait(self, rid):
        """Blocks until done; returns the latency in ms (raises GrglError on failure)."""
        return self._finish(rid, lib().grgl_hip_server_wait)

    def poll(self, rid):
        """Latency in ms when done, None while queued or running."""
        return self._finish(rid, lib().grgl_hip_server_poll)

    def cancel(self, rid):
        check(lib().grgl_hip_server_cancel(self.handle, rid))

    def stats(self, reset=False):
        s = ServerStats()
        check(lib().grgl_hip_server_stats(self.handle, ctypes.byref(s), int(reset)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.handle:
            check(lib().grgl_hip_server_destroy(self.handle))
            self.handle = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
