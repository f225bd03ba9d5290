"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes wrappers for the checker libraries:
  * oracle/_build/libgrg_oracle.so — the plain-C restatement of the reference's CPU
    matmul (grg_oracle.c, following include/grgl/grg.h:1342-1517);
  * oracle/_ref/libgrgl_ref.so     — the reference's own matmul compiled from its
    sources by oracle/Makefile (present only where the reference could be built).

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this package, and
only as the checker or the reported CPU baseline; the product (grgl_amd) never does.
Both expose pygrgl.matmul semantics: a zero-initialised output accumulated into.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "_build", "libgrg_oracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libgrgl_ref.so")

DOWN, UP = 1, 2
INIT_ZERO, INIT_VECTOR, INIT_MATRIX, INIT_XTX = 0, 1, 2, 3
INVALID_NODE = 0x7FFFFFFF
COAL_COUNT_NOT_SET = 0xFFFFFFFF

_DTYPES = {
    np.dtype(np.float64): 0,
    np.dtype(np.float32): 1,
    np.dtype(np.int64): 2,
    np.dtype(np.int32): 3,
    np.dtype(np.int16): 4,
    np.dtype(np.int8): 5,
    np.dtype(np.uint64): 6,
    np.dtype(np.uint32): 7,
    np.dtype(np.uint16): 8,
    np.dtype(np.uint8): 9,
    np.dtype(np.bool_): 10,
}


class OracleMisuse(RuntimeError):
    """The reference's ApiMisuseFailure."""


class _Graph(ctypes.Structure):
    _fields_ = [
        ("num_samples", ctypes.c_uint32),
        ("num_nodes", ctypes.c_uint32),
        ("ploidy", ctypes.c_uint32),
        ("num_mutations", ctypes.c_uint32),
        ("down_off", ctypes.c_void_p),
        ("down_child", ctypes.c_void_p),
        ("mut_node", ctypes.c_void_p),
        ("mut_miss", ctypes.c_void_p),
        ("coal", ctypes.c_void_p),
    ]


def build(reference=None):
    """Build the C restatement (and the reference library when its sources exist)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    ref = reference or "/root/reference"
    if os.path.isfile(os.path.join(ref, "include", "grgl", "grg.h")):
        subprocess.run(["make", "-s", "-C", _HERE, "ref", f"REFERENCE={ref}"], check=True)


_oracle = None
_ref = None


def _load_oracle():
    global _oracle
    if _oracle is None:
        L = ctypes.CDLL(ORACLE_SO)
        vp, sz, ci = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        L.oracle_matmul.restype = ci
        L.oracle_matmul.argtypes = [ctypes.POINTER(_Graph), ci, vp, sz, sz, ci, vp, sz, ci, ci, vp, ci, vp]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_encode.restype = vp
        L.oracle_encode.argtypes = [ctypes.POINTER(_Graph)]
        L.oracle_encoded_free.argtypes = [vp]
        L.oracle_matmul_faithful.restype = ci
        L.oracle_matmul_faithful.argtypes = [vp, ci, vp, sz, ci, vp]
        _oracle = L
    return _oracle


def ref_available():
    return os.path.isfile(REF_SO)


def _load_ref():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_SO)
        vp, sz, ci, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint32
        L.ref_graph_create.restype = vp
        L.ref_graph_create.argtypes = [u32, u32, u32, vp, vp, u32, vp, vp, vp]
        L.ref_graph_destroy.argtypes = [vp]
        L.ref_matmul.restype = ci
        L.ref_matmul.argtypes = [vp, ci, vp, sz, sz, ci, vp, sz, ci, ci, vp, ci, vp]
        L.ref_last_error.restype = ctypes.c_char_p
        L.ref_encode_edges.restype = sz
        L.ref_encode_edges.argtypes = [u32, vp, vp, vp, sz, vp]
        L.ref_mutation_record.restype = ci
        L.ref_mutation_record.argtypes = [ctypes.c_uint64, ctypes.c_char_p, ctypes.c_char_p, u32, vp,
                                          ctypes.c_char_p, sz]
        _ref = L
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data


def ref_mutation_record(position, ref, alt, time=0x3FFFFFFF):
    """(24 record bytes, is_short, allele storage string or None) of the reference's
    Mutation(position, alt, ref, time) as a .grg file stores it."""
    L = _load_ref()
    rec = np.zeros(24, np.uint8)
    buf = ctypes.create_string_buffer(len(ref) + len(alt) + 2)
    short = L.ref_mutation_record(position, ref.encode(), alt.encode(), time, rec.ctypes.data, buf, len(buf))
    return rec, bool(short), (None if short else buf.value.decode())


class OracleGraph:
    """A GRG in the plain arrays of grg_oracle.h (down-edge CSR + mutation map)."""

    def __init__(self, arrays):
        self.num_samples = int(arrays["num_samples"])
        self.ploidy = int(arrays["ploidy"])
        self.off = np.ascontiguousarray(arrays["down_offsets"], dtype=np.uint64)
        self.child = np.ascontiguousarray(arrays["down_children"], dtype=np.uint32)
        self.mut = np.ascontiguousarray(arrays["mutation_nodes"], dtype=np.uint32)
        miss = arrays.get("mutation_miss_nodes")
        coal = arrays.get("coal_counts")
        self.miss = None if miss is None else np.ascontiguousarray(miss, dtype=np.uint32)
        self.coal = None if coal is None else np.ascontiguousarray(coal, dtype=np.uint32)
        self.num_nodes = len(self.off) - 1
        self.num_mutations = len(self.mut)
        self.num_individuals = self.num_samples // self.ploidy
        self._g = _Graph(self.num_samples, self.num_nodes, self.ploidy, self.num_mutations, _ptr(self.off),
                         _ptr(self.child), _ptr(self.mut), _ptr(self.miss), _ptr(self.coal))
        self._ref_handle = None

    def _out_cols(self, direction, emit_all, by_individual):
        if emit_all:
            return self.num_nodes
        if direction == DOWN:
            return self.num_individuals if by_individual else self.num_samples
        return self.num_mutations

    @staticmethod
    def _init_args(init, dtype):
        if init is None:
            return INIT_ZERO, None
        if isinstance(init, str):
            if init != "xtx":
                raise OracleMisuse("Unexpected init value: " + init)
            return INIT_XTX, None
        arr = np.ascontiguousarray(init, dtype=dtype)
        return (INIT_VECTOR if arr.ndim == 1 else INIT_MATRIX), arr

    def matmul(self, inp, direction, emit_all_nodes=False, by_individual=False, init=None, miss=None):
        """oracle_matmul with pygrgl.matmul semantics; `miss` (UP) is accumulated in place."""
        inp = np.ascontiguousarray(inp)
        rows, cols = inp.shape
        out = np.zeros((rows, self._out_cols(direction, emit_all_nodes, by_individual)), dtype=inp.dtype)
        mode, init_arr = self._init_args(init, inp.dtype)
        rc = _load_oracle().oracle_matmul(ctypes.byref(self._g), _DTYPES[inp.dtype], _ptr(inp), cols, rows, direction,
                                          _ptr(out), out.size, int(emit_all_nodes), int(by_individual),
                                          _ptr(init_arr), mode, _ptr(miss))
        if rc != 0:
            raise OracleMisuse(_load_oracle().oracle_last_error().decode())
        return out

    def faithful_matmul(self, inp, direction):
        """Per-node varint decode into fresh buffers (timing mode); f64/f32/i64/i32."""
        inp = np.ascontiguousarray(inp)
        rows, _ = inp.shape
        out = np.zeros((rows, self._out_cols(direction, False, False)), dtype=inp.dtype)
        L = _load_oracle()
        eg = L.oracle_encode(ctypes.byref(self._g))
        try:
            rc = L.oracle_matmul_faithful(eg, _DTYPES[inp.dtype], _ptr(inp), rows, direction, _ptr(out))
        finally:
            L.oracle_encoded_free(eg)
        if rc != 0:
            raise OracleMisuse(L.oracle_last_error().decode())
        return out

    # --- the reference's own matmul (oracle/_ref) --------------------------------------
    def ref_handle(self):
        if self._ref_handle is None:
            L = _load_ref()
            h = L.ref_graph_create(self.num_samples, self.num_nodes, self.ploidy, _ptr(self.off), _ptr(self.child),
                                   self.num_mutations, _ptr(self.mut), _ptr(self.miss), _ptr(self.coal))
            if not h:
                raise RuntimeError("reference graph creation failed: " + L.ref_last_error().decode())
            self._ref_handle = h
        return self._ref_handle

    def ref_matmul(self, inp, direction, emit_all_nodes=False, by_individual=False, init=None, miss=None):
        """The reference's GRG::matrixMultiplication on the same graph (pygrgl semantics)."""
        inp = np.ascontiguousarray(inp)
        rows, cols = inp.shape
        out = np.zeros((rows, self._out_cols(direction, emit_all_nodes, by_individual)), dtype=inp.dtype)
        mode, init_arr = self._init_args(init, inp.dtype)
        L = _load_ref()
        rc = L.ref_matmul(self.ref_handle(), _DTYPES[inp.dtype], _ptr(inp), cols, rows, direction, _ptr(out), out.size,
                          int(emit_all_nodes), int(by_individual), _ptr(init_arr), mode, _ptr(miss))
        if rc != 0:
            raise OracleMisuse(L.ref_last_error().decode())
        return out

    def ref_edge_section(self):
        """(edge bytes, u64 node offsets) of the .grg edge section, encoded by the reference's
        own CSR storage."""
        L = _load_ref()
        n = L.ref_encode_edges(self.num_nodes, _ptr(self.off), _ptr(self.child), None, 0, None)
        buf = np.zeros(n, np.uint8)
        idx = np.zeros(self.num_nodes + 1, np.uint64)
        L.ref_encode_edges(self.num_nodes, _ptr(self.off), _ptr(self.child), _ptr(buf), n, _ptr(idx))
        return buf, idx

    def __del__(self):
        if self._ref_handle is not None and _ref is not None:
            _ref.ref_graph_destroy(self._ref_handle)
            self._ref_handle = None
