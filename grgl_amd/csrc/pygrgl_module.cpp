// _grgl_amd: pybind11 bindings with the pygrgl surface for graph matrix multiplication.
//
// Mirrors src/python/_grgl.cpp of the reference for this path: matmul (:187-241, :1159-1230),
// dispatchMult (:100-185), dot_product (:67-98, :1141-1157), TraversalDirection (:750-757),
// the GRG properties callers use (:354-460, :499-560, :656-710) and the module constants
// (:1266-1269).  All arithmetic runs on the GPU through the C ABI in grgl_hip.h; there is
// no CPU fallback: without a usable device every matmul raises RuntimeError.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>

#include "grgl_hip.h"
#include "construct.h"
#include "fast_build.h"
#include "host_graph.h"
#include "grouping.h"
#include "levelize.h"
#include "stacking.h"

namespace py = pybind11;
using namespace grgl_amd;

namespace {

constexpr size_t kNoUpEdges = std::numeric_limits<size_t>::max(); // include/grgl/grg.h:50

#define API_CHECK(cond, msg)                                                                           \
    do {                                                                                               \
        if (!(cond)) {                                                                                 \
            std::ostringstream _ss;                                                                    \
            _ss << msg;                                                                                \
            throw std::runtime_error(_ss.str());                                                       \
        }                                                                                              \
    } while (0)

void checkStatus(grgl_status st) {
    if (st != GRGL_OK) throw std::runtime_error(grgl_hip_last_error());
}

int g_device = 0;

// Traversal direction with the reference's numeric values (include/grgl/visitor.h:26).
enum class Direction : int { DOWN = 1, UP = 2 };

struct PyGRG {
    HostGrg h;
    std::mutex mu;
    std::mutex syncMu;
    std::map<int, grgl_graph*> dev;
    bool upBuilt = false;
    std::vector<uint64_t> upOff;
    std::vector<uint32_t> upPar;
    std::vector<std::pair<uint32_t, uint32_t>> nodeMut; // (node, mutId) sorted, invalid first

    virtual ~PyGRG() {
        for (auto& kv : dev) grgl_hip_graph_destroy(kv.second);
    }

    // Brings `h` (and the derived caches) up to date; mutable graphs override it.
    virtual void sync() {}
    HostGrg& H() {
        sync();
        return h;
    }

    grgl_graph* deviceGraph() {
        sync();
        std::lock_guard<std::mutex> lk(mu);
        auto it = dev.find(g_device);
        if (it != dev.end()) return it->second;
        grgl_graph_desc d{};
        d.num_samples = h.numSamples;
        d.num_nodes = h.numNodes();
        d.ploidy = h.ploidy;
        d.num_mutations = h.numMutations();
        d.down_offsets = h.downOff.data();
        d.down_children = h.downChild.data();
        d.mutation_nodes = h.mutNode.data();
        d.mutation_miss_nodes = h.mutMiss.empty() ? nullptr : h.mutMiss.data();
        d.coal_counts = h.coal.empty() ? nullptr : h.coal.data();
        grgl_graph* g = nullptr;
        checkStatus(grgl_hip_graph_create(&d, g_device, &g));
        dev[g_device] = g;
        return g;
    }

    void buildUp() {
        sync();
        std::lock_guard<std::mutex> lk(mu);
        if (upBuilt) return;
        const uint32_t V = h.numNodes();
        upOff.assign(V + 1, 0);
        for (uint32_t c : h.downChild) upOff[c + 1]++;
        for (uint32_t v = 0; v < V; v++) upOff[v + 1] += upOff[v];
        upPar.resize(h.numEdges());
        std::vector<uint64_t> cur(upOff.begin(), upOff.end() - 1);
        for (uint32_t p = 0; p < V; p++)
            for (uint64_t e = h.downOff[p]; e < h.downOff[p + 1]; e++) upPar[cur[h.downChild[e]]++] = p;
        nodeMut.clear();
        for (uint32_t m = 0; m < h.numMutations(); m++) nodeMut.push_back({h.mutNode[m], m});
        std::sort(nodeMut.begin(), nodeMut.end(), [](const auto& a, const auto& b) {
            const bool ai = a.first == kInvalidNode, bi = b.first == kInvalidNode;
            if (ai != bi) return ai;
            return a < b;
        });
        upBuilt = true;
    }
};
using PyGRGPtr = std::shared_ptr<PyGRG>;

// MutableGRG (src/python/_grgl.cpp:759-925 of the reference): the growing graph lives in a
// MutableGraph; every read through the GRG interface first refreezes it into `h` when it
// changed, dropping device copies of the old version.
struct PyMutableGRG : PyGRG {
    MutableGraph mg;
    uint64_t syncedVersion = ~0ULL;
    PyMutableGRG(uint32_t numSamples, uint32_t ploidy, bool phased) : mg(numSamples, ploidy, phased) {}
    void sync() override {
        std::lock_guard<std::mutex> sl(syncMu);
        if (mg.version() == syncedVersion) return;
        HostGrg fresh = mg.freeze();
        std::lock_guard<std::mutex> lk(mu);
        for (auto& kv : dev) grgl_hip_graph_destroy(kv.second);
        dev.clear();
        upBuilt = false;
        h = std::move(fresh);
        syncedVersion = mg.version();
    }
};
using PyMutableGRGPtr = std::shared_ptr<PyMutableGRG>;

// A mutation's public fields (include/grgl/mutation.h): position, ALT, REF, time.
struct PyMutation {
    uint64_t position = 0;
    std::string allele;
    std::string ref_allele;
    uint32_t time = 0;
};

grgl_dtype dtypeOf(py::handle input, bool& ok) {
    ok = true;
    if (py::isinstance<py::array_t<double>>(input)) return GRGL_F64;
    if (py::isinstance<py::array_t<float>>(input)) return GRGL_F32;
    if (py::isinstance<py::array_t<int64_t>>(input)) return GRGL_I64;
    if (py::isinstance<py::array_t<int32_t>>(input)) return GRGL_I32;
    if (py::isinstance<py::array_t<int16_t>>(input)) return GRGL_I16;
    if (py::isinstance<py::array_t<int8_t>>(input)) return GRGL_I8;
    if (py::isinstance<py::array_t<uint64_t>>(input)) return GRGL_U64;
    if (py::isinstance<py::array_t<uint32_t>>(input)) return GRGL_U32;
    if (py::isinstance<py::array_t<uint16_t>>(input)) return GRGL_U16;
    if (py::isinstance<py::array_t<uint8_t>>(input)) return GRGL_U8;
    if (py::isinstance<py::array_t<bool>>(input)) return GRGL_BOOL;
    ok = false;
    return GRGL_F64;
}

py::dtype numpyDtype(grgl_dtype dt) {
    switch (dt) {
    case GRGL_F64: return py::dtype::of<double>();
    case GRGL_F32: return py::dtype::of<float>();
    case GRGL_I64: return py::dtype::of<int64_t>();
    case GRGL_I32: return py::dtype::of<int32_t>();
    case GRGL_I16: return py::dtype::of<int16_t>();
    case GRGL_I8: return py::dtype::of<int8_t>();
    case GRGL_U64: return py::dtype::of<uint64_t>();
    case GRGL_U32: return py::dtype::of<uint32_t>();
    case GRGL_U16: return py::dtype::of<uint16_t>();
    case GRGL_U8: return py::dtype::of<uint8_t>();
    default: return py::dtype::of<bool>();
    }
}

bool sameDtype(py::handle h, grgl_dtype dt) {
    bool ok = false;
    const grgl_dtype got = dtypeOf(h, ok);
    return ok && got == dt;
}

// dispatchMult (src/python/_grgl.cpp:100-185): zeroed result, init/miss parsing, GIL released.
// Page-locked host blocks that back large results, so the device-to-host copy of the output
// runs at link speed instead of through the runtime's pageable staging (UP K = 16 on a
// 1M-mutation graph copies 64 MB back).  A result's block returns to the pool when numpy
// frees the array; the pool keeps at most kMaxFree bytes of idle blocks, and at most
// kMaxLive bytes are handed out (beyond that results are ordinary numpy memory).  The pool
// is never destroyed (blocks may outlive the module's statics at interpreter exit).
class PinnedPool {
public:
    static PinnedPool& get() {
        static PinnedPool* pool = new PinnedPool();
        return *pool;
    }
    static constexpr size_t kMinBytes = 1ULL << 20;
    void* acquire(size_t bytes) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            freeBytes_ -= it->first;
            live_ += it->first;
            free_.erase(it);
            return p;
        }
        if (live_ + bytes > kMaxLive) return nullptr;
        void* p = nullptr;
        if (grgl_hip_host_alloc(bytes, &p) != GRGL_OK) return nullptr;
        size_[p] = bytes;
        live_ += bytes;
        return p;
    }
    void release(void* p) {
        std::lock_guard<std::mutex> lk(mu_);
        const size_t b = size_[p];
        live_ -= b;
        if (freeBytes_ + b <= kMaxFree) {
            free_.emplace(b, p);
            freeBytes_ += b;
        } else {
            size_.erase(p);
            (void)grgl_hip_host_free(p);
        }
    }

private:
    static constexpr size_t kMaxFree = 2ULL << 30, kMaxLive = 16ULL << 30;
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    std::map<void*, size_t> size_;
    size_t freeBytes_ = 0, live_ = 0;
};

// A C-contiguous array for a result; large ones live in the pinned pool.
py::array resultArray(grgl_dtype dt, std::vector<py::ssize_t> shape, size_t elems) {
    const py::dtype d = numpyDtype(dt);
    const size_t bytes = elems * (size_t)d.itemsize();
    if (bytes >= PinnedPool::kMinBytes && grgl_hip_device_count() > 0) {
        if (void* p = PinnedPool::get().acquire(bytes)) {
            py::capsule owner(p, [](void* q) { PinnedPool::get().release(q); });
            return py::array(d, shape, p, owner);
        }
    }
    return py::array(d, shape);
}

py::array dispatchMult(PyGRG& grg, grgl_dtype dt, const py::array& arr, size_t rows, size_t inCols, size_t outCols,
                       Direction direction, bool emitAllNodes, bool byIndividual, py::handle init, py::handle miss,
                       uint32_t extraFlags) {
    const size_t outSize = rows * outCols;
    // The reference zero-fills the result and accumulates into it (_grgl.cpp:112-114).  The
    // device path writes every output element as 0 + value, which is the same, so the result
    // is filled only where the call accumulates anyway: UP with `miss`, whose missingness
    // output is accumulated into the caller's array.
    const bool accumulate = !miss.is_none() && direction == Direction::UP;
    py::array result = resultArray(dt, {(py::ssize_t)rows, (py::ssize_t)outCols}, outSize);
    if (accumulate) std::memset(result.mutable_data(), 0, outSize * result.itemsize());

    py::array initArr;
    const void* initPtr = nullptr;
    grgl_init_mode initMode = GRGL_INIT_ZERO;
    if (init.is_none()) {
        ;
    } else if (py::isinstance<py::str>(init)) {
        const std::string strval = init.cast<std::string>();
        API_CHECK(strval == "xtx", "Unexpected init value: " << strval);
        initMode = GRGL_INIT_XTX;
    } else {
        initArr = py::array::ensure(init, py::array::c_style | py::array::forcecast);
        API_CHECK(sameDtype(init, dt), "The init matrix must match the dtype of the input matrix. Got: "
                                           << std::string(py::str(initArr.dtype())));
        initPtr = initArr.data();
        if (initArr.ndim() == 1) {
            API_CHECK((size_t)initArr.shape(0) == rows,
                      "If init has a single dimension, it must match the number of rows in the input matrix");
            initMode = GRGL_INIT_VECTOR;
        }
        if (initArr.ndim() == 2) {
            API_CHECK((size_t)initArr.shape(0) == rows && (size_t)initArr.shape(1) == grg.H().numNodes(),
                      "If init is a matrix, it must match the dimensions ROW x NODES");
            initMode = GRGL_INIT_MATRIX;
        }
    }

    py::array missArr;
    void* missPtr = nullptr;
    if (!miss.is_none()) {
        API_CHECK(!emitAllNodes, "The \"miss\" parameter cannot be mixed with the \"emit_all_nodes\" parameter");
        missArr = py::array::ensure(miss, py::array::c_style | py::array::forcecast);
        API_CHECK(sameDtype(miss, dt), "The \"miss\" input must be a matrix with dtype matching the input matrix. Got: "
                                           << std::string(py::str(missArr.dtype())));
        API_CHECK(missArr.ndim() == 2, "\"miss\" must be a two-dimension numpy array (matrix). ndim=" << missArr.ndim());
        const size_t missRows = missArr.shape(0), missCols = missArr.shape(1);
        API_CHECK(missRows == rows, "\"miss\" has " << missRows << " rows, but must match the input and output matrices ("
                                                    << rows << ")");
        if (direction == Direction::DOWN) {
            API_CHECK(missCols == inCols,
                      "The \"miss\" matrix must match the number of columns in the input matrix. Got: " << missCols);
        } else {
            API_CHECK(missCols == outCols,
                      "The \"miss\" matrix must match the number of columns in the output matrix. Got: " << missCols);
        }
        missPtr = missArr.mutable_data();
    }

    grgl_graph* g = grg.deviceGraph();
    const void* in = arr.data();
    void* out = result.mutable_data();
    grgl_status st;
    {
        py::gil_scoped_release release;
        st = grgl_hip_matmul(g, dt, in, inCols, rows, (grgl_direction)direction, out, outSize, emitAllNodes ? 1 : 0,
                             byIndividual ? 1 : 0, initPtr, initMode, missPtr,
                             GRGL_FLAG_HOST_PTRS | (accumulate ? GRGL_FLAG_ACCUMULATE : 0U) | extraFlags, nullptr);
    }
    checkStatus(st);
    return result;
}

// matMul (src/python/_grgl.cpp:187-241).
py::array matMul(PyGRGPtr& grg, py::handle input, Direction direction, bool emitAllNodes, bool byIndividual,
                 py::handle init, py::handle miss, bool orderPreserving) {
    py::array arr = py::array::ensure(input, py::array::c_style | py::array::forcecast);
    API_CHECK(arr, "matmul() requires a numpy array as input.");
    API_CHECK(arr.ndim() == 2, "matmul() only support two dimensional numpy arrays as input.");
    const size_t rows = arr.shape(0), cols = arr.shape(1);
    API_CHECK(rows != 0 && cols != 0, "matmul() requires non-zero dimensions.");
    const size_t numSamples = byIndividual ? grg->H().numSamples / grg->H().ploidy : grg->H().numSamples;
    const size_t outCols =
        emitAllNodes ? grg->H().numNodes() : (direction == Direction::DOWN ? numSamples : grg->H().numMutations());
    bool ok = false;
    const grgl_dtype dt = dtypeOf(input, ok);
    API_CHECK(ok, "Unsupported numpy dtype: " << std::string(py::str(arr.dtype())));
    if (dt == GRGL_BOOL) API_CHECK(init.is_none(), "init=... is not supported with boolean dtype");
    return dispatchMult(*grg, dt, arr, rows, cols, outCols, direction, emitAllNodes, byIndividual, init, miss,
                        orderPreserving ? GRGL_FLAG_ORDER_PRESERVING : 0U);
}

// dotProduct (src/python/_grgl.cpp:67-98).
py::array dotProduct(PyGRGPtr& grg, py::handle input, Direction direction) {
    py::array arr = py::array::ensure(input, py::array::c_style | py::array::forcecast);
    API_CHECK(arr, "dot_product() requires a numpy array as input.");
    API_CHECK(arr.ndim() == 1, "dot_product() only support one-dimensional numpy arrays as input.");
    const size_t cols = arr.shape(0);
    API_CHECK(cols != 0, "dot_product() requires non-empty array input.");
    const size_t outCols = direction == Direction::DOWN ? grg->H().numSamples : grg->H().numMutations();
    bool ok = false;
    const grgl_dtype dt = dtypeOf(input, ok);
    API_CHECK(ok && (dt == GRGL_F64 || dt == GRGL_F32 || dt == GRGL_I64 || dt == GRGL_I32),
              "Unsupported numpy dtype: " << std::string(py::str(arr.dtype())));
    py::array result(numpyDtype(dt), std::vector<py::ssize_t>{(py::ssize_t)outCols}); // every element written
    grgl_graph* g = grg->deviceGraph();
    const void* in = arr.data();
    void* out = result.mutable_data();
    grgl_status st;
    {
        py::gil_scoped_release release;
        st = grgl_hip_matmul(g, dt, in, cols, 1, (grgl_direction)direction, out, outCols, 0, 0, nullptr, GRGL_INIT_ZERO,
                             nullptr, GRGL_FLAG_HOST_PTRS, nullptr);
    }
    checkStatus(st);
    return result;
}

template <typename T> std::vector<T> toVec(py::handle h, const char* what) {
    py::array_t<T, py::array::c_style | py::array::forcecast> a = py::array_t<T, py::array::c_style | py::array::forcecast>::ensure(h);
    API_CHECK(a, what << " must be array-like");
    API_CHECK(a.ndim() == 1, what << " must be one-dimensional");
    return std::vector<T>(a.data(), a.data() + a.size());
}

template <typename T> py::array_t<T> toArray(const std::vector<T>& v) {
    py::array_t<T> a((py::ssize_t)v.size());
    if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(T));
    return a;
}

PyGRGPtr fromArrays(uint32_t numSamples, uint32_t ploidy, py::handle downOffsets, py::handle downChildren,
                    py::handle mutationNodes, py::handle mutationMiss, py::handle coalCounts, py::handle positions,
                    py::handle refAlleles, py::handle altAlleles, py::handle populations, py::handle populationIds,
                    py::handle individualIds) {
    auto g = std::make_shared<PyGRG>();
    g->H().numSamples = numSamples;
    g->H().ploidy = ploidy;
    g->H().downOff = toVec<uint64_t>(downOffsets, "down_offsets");
    g->H().downChild = toVec<uint32_t>(downChildren, "down_children");
    g->H().mutNode = toVec<uint32_t>(mutationNodes, "mutation_nodes");
    if (!mutationMiss.is_none()) g->H().mutMiss = toVec<uint32_t>(mutationMiss, "mutation_miss_nodes");
    if (!coalCounts.is_none()) g->H().coal = toVec<uint32_t>(coalCounts, "coal_counts");
    if (!positions.is_none()) g->H().mutPos = toVec<double>(positions, "positions");
    API_CHECK(!g->H().downOff.empty(), "down_offsets must have num_nodes + 1 entries");
    if (!refAlleles.is_none() || !altAlleles.is_none()) {
        const auto refs = refAlleles.cast<std::vector<std::string>>();
        const auto alts = altAlleles.cast<std::vector<std::string>>();
        API_CHECK(refs.size() == g->H().numMutations() && alts.size() == g->H().numMutations(),
                  "ref_alleles and alt_alleles need one entry per mutation");
        for (size_t i = 0; i < refs.size(); i++) {
            API_CHECK(refs[i].size() <= 0xffff, "REF allele length exceeded maximum of 65535");
            g->H().mutAlleles.push_back(refs[i] + alts[i]);
            g->H().mutAltIndex.push_back((uint16_t)refs[i].size());
        }
    }
    if (!populations.is_none()) g->H().populations = populations.cast<std::vector<std::string>>();
    if (!populationIds.is_none()) g->H().popIds = toVec<uint16_t>(populationIds, "population_ids");
    if (!individualIds.is_none()) {
        g->H().individualIds = individualIds.cast<std::vector<std::string>>();
        API_CHECK(g->H().individualIds.size() == numSamples / ploidy, "individual_ids needs one entry per individual");
    }
    validateHostGrg(g->H());
    return g;
}

} // namespace

PYBIND11_MODULE(_grgl_amd, m) {
    m.doc() = "MI355X-native GRG matrix multiplication (pygrgl-compatible surface)";

    py::enum_<Direction>(m, "TraversalDirection")
        .value("DOWN", Direction::DOWN, "Traverse the graph \"down\" edges.")
        .value("UP", Direction::UP, "Traverse the graph via \"up\" edges.")
        .export_values();

    py::class_<PyMutation>(m, "Mutation")
        .def(py::init([](double position, const std::string& allele, const std::string& refAllele, double time) {
                 PyMutation x;
                 x.position = (uint64_t)position;
                 x.allele = allele;
                 x.ref_allele = refAllele;
                 x.time = time < 0 ? 0x3FFFFFFFU : (uint32_t)time; // MUT_TIME_NOT_SET, mutation.h:44
                 return x;
             }),
             py::arg("position"), py::arg("allele"), py::arg("ref_allele") = "", py::arg("time") = -1.0,
             "A mutation to add to a MutableGRG (src/python/_grgl.cpp:304-312 of the reference).")
        .def_readonly("position", &PyMutation::position)
        .def_readonly("allele", &PyMutation::allele)
        .def_readonly("ref_allele", &PyMutation::ref_allele)
        .def_readonly("time", &PyMutation::time)
        .def("__repr__", [](const PyMutation& x) {
            return "Mutation(" + std::to_string(x.position) + ", " + x.ref_allele + ">" + x.allele + ")";
        });

    py::class_<PyGRG, PyGRGPtr>(m, "GRG")
        .def_property_readonly("num_samples", [](PyGRG& g) { return g.H().numSamples; })
        .def_property_readonly("num_individuals", [](PyGRG& g) { return g.H().numSamples / g.H().ploidy; })
        .def_property_readonly("ploidy", [](PyGRG& g) { return g.H().ploidy; })
        .def_property_readonly("shape",
                               [](PyGRG& g) { return std::make_pair(g.H().numSamples / g.H().ploidy, g.H().numMutations()); })
        .def_property_readonly("num_nodes", [](PyGRG& g) { return g.H().numNodes(); })
        .def_property_readonly("num_edges", [](PyGRG& g) { return g.H().numEdges(); })
        .def_property_readonly("is_phased", [](PyGRG& g) { return g.H().phased; })
        .def_property_readonly("num_mutations", [](PyGRG& g) { return g.H().numMutations(); })
        .def_property_readonly("has_missing_data", [](PyGRG& g) { return !g.H().mutMiss.empty(); })
        .def_property_readonly("has_individual_coals", [](PyGRG& g) { return g.H().hasCoals(); })
        .def_property_readonly("nodes_are_ordered", [](PyGRG&) { return true; })
        .def_property_readonly("is_mutable", [](PyGRG&) { return false; })
        .def("is_sample", [](PyGRG& g, uint32_t n) { return n < g.H().numSamples; }, py::arg("node_id"))
        .def("num_down_edges",
             [](PyGRG& g, uint32_t n) {
                 API_CHECK(n < g.H().numNodes(), "Index out of bounds (CSR data)");
                 return (size_t)(g.H().downOff[n + 1] - g.H().downOff[n]);
             },
             py::arg("node_id"))
        .def("get_down_edges",
             [](PyGRG& g, uint32_t n) {
                 API_CHECK(n < g.H().numNodes(), "Index out of bounds (CSR data)");
                 return std::vector<uint32_t>(g.H().downChild.begin() + g.H().downOff[n],
                                              g.H().downChild.begin() + g.H().downOff[n + 1]);
             },
             py::arg("node_id"))
        .def("num_up_edges",
             [](PyGRG& g, uint32_t n) {
                 g.buildUp();
                 API_CHECK(n < g.H().numNodes(), "Index out of bounds (CSR data)");
                 return (size_t)(g.upOff[n + 1] - g.upOff[n]);
             },
             py::arg("node_id"))
        .def("get_up_edges",
             [](PyGRG& g, uint32_t n) {
                 g.buildUp();
                 API_CHECK(n < g.H().numNodes(), "Index out of bounds (CSR data)");
                 return std::vector<uint32_t>(g.upPar.begin() + g.upOff[n], g.upPar.begin() + g.upOff[n + 1]);
             },
             py::arg("node_id"))
        .def("get_mutation_node_pairs",
             [](PyGRG& g) {
                 std::vector<std::pair<uint32_t, uint32_t>> r;
                 for (uint32_t m = 0; m < g.H().numMutations(); m++) r.push_back({m, g.H().mutNode[m]});
                 return r;
             })
        .def("get_mutation_node_miss",
             [](PyGRG& g) {
                 std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> r;
                 for (uint32_t m = 0; m < g.H().numMutations(); m++)
                     r.push_back({m, g.H().mutNode[m], g.H().mutMiss.empty() ? kInvalidNode : g.H().mutMiss[m]});
                 return r;
             })
        .def("get_mutations_for_node",
             [](PyGRG& g, uint32_t n) {
                 g.buildUp();
                 std::vector<uint32_t> r;
                 auto it = std::lower_bound(g.nodeMut.begin(), g.nodeMut.end(), std::make_pair(n, 0U),
                                            [](const auto& a, const auto& b) {
                                                const bool ai = a.first == kInvalidNode, bi = b.first == kInvalidNode;
                                                if (ai != bi) return ai;
                                                return a < b;
                                            });
                 for (; it != g.nodeMut.end() && it->first == n; ++it) r.push_back(it->second);
                 return r;
             },
             py::arg("node_id"))
        .def("get_num_individual_coals",
             [](PyGRG& g, uint32_t n) {
                 if (n < g.H().numSamples) return 0U;
                 if (g.H().coal.empty() || n >= g.H().numNodes()) return kCoalNotSet;
                 return g.H().coal[n];
             },
             py::arg("node_id"))
        .def("get_mutation_position", [](PyGRG& g, uint32_t m) {
            API_CHECK(m < g.H().numMutations(), "Invalid MutationID: " << m);
            return g.H().mutPos.empty() ? (double)m : g.H().mutPos[m];
        })
        .def("get_mutation_by_id",
             [](PyGRG& g, uint32_t m) {
                 API_CHECK(m < g.H().numMutations(), "Invalid MutationID: " << m);
                 PyMutation r;
                 r.position = g.H().mutPos.empty() ? (uint64_t)m : (uint64_t)g.H().mutPos[m];
                 const std::string s = g.H().mutAlleles.empty() ? std::string("AG") : g.H().mutAlleles[m];
                 const size_t alt = g.H().mutAltIndex.empty() ? 1 : std::min<size_t>(g.H().mutAltIndex[m], s.size());
                 r.ref_allele = s.substr(0, alt);
                 r.allele = s.substr(alt);
                 r.time = g.H().mutTime.empty() ? 0x3FFFFFFFU : g.H().mutTime[m];
                 return r;
             },
             py::arg("mut_id"))
        .def_property_readonly("specified_bp_range",
                               [](PyGRG& g) { return std::make_pair(g.H().rangeStart, g.H().rangeEnd); })
        .def_property_readonly("num_populations", [](PyGRG& g) { return g.H().populations.size(); })
        .def("get_populations", [](PyGRG& g) { return g.H().populations; })
        .def_property_readonly("has_individual_ids", [](PyGRG& g) { return !g.H().individualIds.empty(); })
        .def("get_individual_id",
             [](PyGRG& g, uint32_t i) {
                 API_CHECK(i < g.H().individualIds.size(), "Invalid individual index " << i);
                 return g.H().individualIds[i];
             },
             py::arg("individual_index"))
        .def("arrays",
             [](PyGRG& g) {
                 py::dict d;
                 d["num_samples"] = g.H().numSamples;
                 d["ploidy"] = g.H().ploidy;
                 d["down_offsets"] = toArray(g.H().downOff);
                 d["down_children"] = toArray(g.H().downChild);
                 d["mutation_nodes"] = toArray(g.H().mutNode);
                 d["mutation_miss_nodes"] = g.H().mutMiss.empty() ? py::object(py::none()) : py::object(toArray(g.H().mutMiss));
                 d["coal_counts"] = g.H().coal.empty() ? py::object(py::none()) : py::object(toArray(g.H().coal));
                 d["positions"] = toArray(g.H().mutPos);
                 return d;
             },
             "The graph as plain arrays (down-edge CSR, mutation map, coalescence counts).")
        .def("mutation_carriers",
             [](PyGRG& g) {
                 CarrierParts cp;
                 {
                     py::gil_scoped_release release;
                     cp = mutationCarriers(g.H());
                 }
                 py::array_t<uint32_t> samples((py::ssize_t)cp.off.back());
                 uint32_t* dst = samples.mutable_data();
                 for (auto& part : cp.parts) { // one part at a time: the peak stays near one copy
                     if (!part.empty()) std::memcpy(dst, part.data(), part.size() * sizeof(uint32_t));
                     dst += part.size();
                     std::vector<uint32_t>().swap(part);
                 }
                 return py::make_tuple(toArray(cp.off), samples);
             },
             "The samples carrying each mutation as CSR arrays (offsets of length num_mutations+1, sorted "
             "sample IDs): the sparse genotype matrix the graph encodes.")
        .def("graph_stats",
             [](PyGRG& g) {
                 const GraphStats st = computeGraphStats(g.H());
                 py::dict d;
                 d["num_nodes"] = g.H().numNodes();
                 d["num_edges"] = st.numEdges;
                 d["up_levels"] = st.upLevels;
                 d["max_fan_in"] = st.maxFanIn;
                 d["fan_in_histogram"] = st.fanInHistogram;
                 return d;
             })
        .def("device_info",
             [](PyGRG& g) {
                 grgl_graph_info info{};
                 checkStatus(grgl_hip_graph_info(g.deviceGraph(), &info));
                 py::dict d;
                 d["num_nodes"] = info.num_nodes;
                 d["num_edges"] = info.num_edges;
                 d["up_levels"] = info.up_levels;
                 d["down_levels"] = info.down_levels;
                 d["max_fanin_up"] = info.max_fanin_up;
                 d["max_fanin_down"] = info.max_fanin_down;
                 d["up_gather_entries"] = info.up_gather_entries;
                 d["down_gather_entries"] = info.down_gather_entries;
                 d["device_bytes"] = info.device_bytes;
                 d["device"] = info.device;
                 return d;
             })
        .def("device_handle", [](PyGRG& g) { return (uintptr_t)g.deviceGraph(); },
             "Address of the device-resident graph (grgl_graph*) for C-ABI callers.")
        .def("__str__",
             [](PyGRG& g) {
                 std::ostringstream ss;
                 ss << "<[grgl_amd.GRG] mutations=" << g.H().numMutations() << ", individuals=" << g.H().numSamples / g.H().ploidy
                    << ", ploidy=" << g.H().ploidy << ", mutable=No>";
                 return ss.str();
             })
        .def("__repr__", [](PyGRG& g) {
            std::ostringstream ss;
            ss << "<[grgl_amd.GRG] mutations=" << g.H().numMutations() << ", individuals=" << g.H().numSamples / g.H().ploidy
               << ", ploidy=" << g.H().ploidy << ", mutable=No>";
            return ss.str();
        });

    py::class_<MappingStats>(m, "MutationMappingStats")
        .def_readonly("total_mutations", &MappingStats::totalMutations)
        .def_readonly("empty_mutations", &MappingStats::emptyMutations)
        .def_readonly("mutations_with_one_sample", &MappingStats::mutationsWithOneSample)
        .def_readonly("mutations_with_no_candidates", &MappingStats::mutationsWithNoCandidates)
        .def_readonly("reused_nodes", &MappingStats::reusedNodes)
        .def_readonly("reused_node_coverage", &MappingStats::reusedNodeCoverage)
        .def_readonly("reused_exactly", &MappingStats::reusedExactly)
        .def_readonly("singleton_sample_edges", &MappingStats::singletonSampleEdges)
        .def_readonly("new_tree_nodes", &MappingStats::newTreeNodes)
        .def_readonly("samples_processed", &MappingStats::samplesProcessed)
        .def_readonly("num_candidates", &MappingStats::numCandidates)
        .def_readonly("reuse_size_bigger_than_hist_max", &MappingStats::reuseSizeBiggerThanHistMax)
        .def_readonly("num_with_singletons", &MappingStats::numWithSingletons)
        .def_readonly("max_singletons", &MappingStats::maxSingletons)
        .def_readonly("reused_mut_nodes", &MappingStats::reusedMutNodes)
        .def_readonly("reuse_size_hist", &MappingStats::reuseSizeHist);

    auto toRecord = [](const PyMutation& x) {
        MutationRecordIn r;
        r.position = (double)x.position;
        r.allele = x.allele;
        r.refAllele = x.ref_allele;
        return r;
    };
    py::class_<PyMutableGRG, PyGRG, PyMutableGRGPtr>(m, "MutableGRG")
        .def(py::init([](uint32_t numSamples, uint32_t ploidy, bool phased, size_t initialNodeCapacity) {
                 (void)initialNodeCapacity;
                 return std::make_shared<PyMutableGRG>(numSamples, ploidy, phased);
             }),
             py::arg("num_samples"), py::arg("ploidy"), py::arg("phased") = true,
             py::arg("initial_node_capacity") = 1024,
             "A GRG that can grow: nodes 0..num_samples-1 are the samples (reference: "
             "src/python/_grgl.cpp:759-775).  New nodes come after every existing node, and edges must go "
             "from a node to one with a smaller ID, so the graph stays ordered.")
        .def(
            "make_node",
            [](PyMutableGRG& g, uint32_t count, bool negative) {
                API_CHECK(!negative, "negative nodes are not supported: new nodes always come last");
                return g.mg.makeNode(count);
            },
            py::arg("count") = 1, py::arg("negative") = false)
        .def(
            "connect", [](PyMutableGRG& g, uint32_t source, uint32_t target) { g.mg.connect(source, target); },
            py::arg("source"), py::arg("target"))
        .def(
            "add_mutation",
            [toRecord](PyMutableGRG& g, const PyMutation& mut, uint32_t node, uint32_t miss) {
                return g.mg.addMutation(toRecord(mut), node, miss);
            },
            py::arg("mutation"), py::arg("node_id"), py::arg("miss_node_id") = kInvalidNode)
        .def(
            "set_num_individual_coals",
            [](PyMutableGRG& g, uint32_t node, uint32_t coals) { g.mg.setNumIndividualCoals(node, coals); },
            py::arg("node_id"), py::arg("num_coals"))
        .def("__repr__", [](PyMutableGRG& g) {
            std::ostringstream ss;
            ss << "<[grgl_amd.GRG] mutations=" << g.mg.numMutations() << ", individuals="
               << g.mg.numSamples() / g.mg.ploidy() << ", ploidy=" << g.mg.ploidy() << ", mutable=Yes>";
            return ss.str();
        });

    m.def(
        "map_mutations",
        [toRecord](PyMutableGRG& g, const std::vector<PyMutation>& mutations,
                   const std::vector<std::vector<uint32_t>>& samples, bool verbose, size_t batchSize,
                   bool sizeOrder) {
            (void)verbose;
            std::vector<MutationRecordIn> recs;
            recs.reserve(mutations.size());
            for (const auto& x : mutations) recs.push_back(toRecord(x));
            py::gil_scoped_release release;
            return g.mg.mapMutations(recs, samples, batchSize, sizeOrder);
        },
        py::arg("grg"), py::arg("mutations"), py::arg("samples"), py::arg("verbose") = false,
        py::arg("mutation_batch_size") = 0, py::kw_only(), py::arg("size_order") = false,
        "Map mutations into a MutableGRG so that the samples below each mutation's node are exactly its "
        "sample list (the contract of the reference's map_mutations, src/python/_grgl.cpp:1029-1062).  "
        "Nodes already in the graph when a batch starts are reused where they fit.  size_order=True (not in "
        "the reference) maps the sets smallest first; with mutation_batch_size=1 this builds a compact graph "
        "from an empty one.  Mutation IDs follow the input order.");

    py::class_<FastBuildStats>(m, "GraphBuildStats")
        .def_readonly("trees", &FastBuildStats::trees)
        .def_readonly("windows", &FastBuildStats::windows)
        .def_readonly("unique_hap_segments", &FastBuildStats::uniqueHapSegments)
        .def_readonly("direct_mapped", &FastBuildStats::directMapped)
        .def_readonly("tree_mapped", &FastBuildStats::treeMapped)
        .def_readonly("unmapped", &FastBuildStats::unmapped)
        .def_readonly("comparisons", &FastBuildStats::comparisons)
        .def_readonly("queries", &FastBuildStats::queries)
        .def_readonly("truncated_queries", &FastBuildStats::truncatedQueries)
        .def_readonly("rebuilds", &FastBuildStats::rebuilds)
        .def_readonly("combined_nodes", &FastBuildStats::combinedNodes);

    m.def(
        "fast_grg_from_samples",
        [toRecord](uint32_t numSamples, uint32_t ploidy, py::object mutations, py::object samples,
                   py::object carrierOffsets, py::object carriers, py::object positions, py::object missing,
                   py::object trees,
                   double lfNoTree, bool noTreeMap, double rebuildProportion, uint32_t threads, bool returnStats,
                   bool phased, uint32_t hapLength) {
            FastBuildParams p;
            p.hapLength = hapLength;
            p.noTreeBelow = lfNoTree;
            p.treeMap = !noTreeMap;
            p.rebuildProportion = rebuildProportion;
            // The --trees option of the reference's builder (src/grgl.cpp:295-311).
            if (py::isinstance<py::str>(trees)) {
                const std::string t = trees.cast<std::string>();
                if (t == "faster1") {
                    p.hapSumMultiplier = 0.75;
                } else if (t == "faster2") {
                    p.hapSumMultiplier = 0.5;
                } else if (t != "optimal") {
                    throw py::value_error("trees must be 'optimal', 'faster1', 'faster2' or a positive integer");
                }
            } else {
                const long long n = trees.cast<long long>();
                API_CHECK(n > 0, "Must provide non-zero trees argument");
                p.treeCount = (uint32_t)n;
            }
            // Carrier lists as CSR (offsets checked here; rows are sorted by the builder if needed).
            std::vector<uint64_t> off(1, 0);
            std::vector<uint32_t> flat;
            py::array_t<uint32_t, py::array::c_style | py::array::forcecast> carArr;
            const uint32_t* carPtr = nullptr;
            if (!samples.is_none()) {
                API_CHECK(carrierOffsets.is_none() && carriers.is_none(), "give samples or carrier arrays, not both");
                for (const auto& c : samples.cast<std::vector<std::vector<uint32_t>>>()) {
                    flat.insert(flat.end(), c.begin(), c.end());
                    off.push_back(flat.size());
                }
                carPtr = flat.data();
            } else {
                API_CHECK(!carrierOffsets.is_none() && !carriers.is_none(),
                          "give samples (a list per mutation) or carrier_offsets and carriers");
                auto offArr = carrierOffsets.cast<py::array_t<int64_t, py::array::c_style | py::array::forcecast>>();
                carArr = carriers.cast<py::array_t<uint32_t, py::array::c_style | py::array::forcecast>>();
                API_CHECK(offArr.ndim() == 1 && offArr.size() >= 1, "carrier_offsets must be a 1-D array of M+1 offsets");
                const int64_t* o = offArr.data();
                API_CHECK(o[0] == 0, "carrier_offsets must start at 0");
                off.resize((size_t)offArr.size());
                for (size_t i = 1; i < (size_t)offArr.size(); i++) {
                    API_CHECK(o[i] >= o[i - 1] && o[i] <= (int64_t)carArr.size(), "bad carrier_offsets");
                    off[i] = (uint64_t)o[i];
                }
                carPtr = carArr.data();
            }
            const size_t M = off.size() - 1;
            std::vector<MutationRecordIn> recs;
            if (!mutations.is_none()) {
                for (const auto& x : mutations.cast<std::vector<PyMutation>>()) recs.push_back(toRecord(x));
            } else {
                std::vector<double> pos;
                if (!positions.is_none()) pos = positions.cast<std::vector<double>>();
                API_CHECK(pos.empty() || pos.size() == M, "positions must have one entry per mutation");
                std::vector<bool> miss;
                if (!missing.is_none()) miss = missing.cast<std::vector<bool>>();
                API_CHECK(miss.empty() || miss.size() == M, "missing must have one entry per mutation");
                recs.resize(M);
                for (size_t i = 0; i < M; i++) {
                    recs[i].position = pos.empty() ? (double)i : pos[i];
                    recs[i].allele = (!miss.empty() && miss[i]) ? "." : "G";
                    recs[i].refAllele = "A";
                }
            }
            API_CHECK(recs.size() == M, "mutations and samples must be the same length");
            auto g = std::make_shared<PyMutableGRG>(numSamples, ploidy, phased);
            FastBuildStats st;
            {
                py::gil_scoped_release release;
                g->mg = fastBuildFromSamples(numSamples, ploidy, recs, off.data(), carPtr, p, st, threads, phased);
            }
            if (returnStats) return py::object(py::make_tuple(g, st));
            return py::object(py::cast(g));
        },
        py::arg("num_samples"), py::arg("ploidy"), py::arg("mutations") = py::none(), py::arg("samples") = py::none(),
        py::kw_only(), py::arg("carrier_offsets") = py::none(), py::arg("carriers") = py::none(),
        py::arg("positions") = py::none(), py::arg("missing") = py::none(), py::arg("trees") = py::str("optimal"),
        py::arg("lf_no_tree") = 0.0,
        py::arg("no_tree_map") = false, py::arg("rebuild_proportion") = 0.10, py::arg("threads") = 0,
        py::arg("return_stats") = false, py::arg("phased") = true, py::arg("hap_length") = 254,
        "Build a GRG from haplotypes, as the reference's graph construction does (fastGRGFromSamples, "
        "src/fast_build.cpp:794-986): the variant rows are cut into trees, each tree clusters the samples by "
        "haplotype-segment similarity and maps its rows onto the roots, and the trees are merged combining equal "
        "nodes.  Input: one carrier list per mutation (samples=...), or CSR arrays (carrier_offsets, carriers), in "
        "genome order.  trees: 'optimal', 'faster1', 'faster2' or a count; lf_no_tree: rows carried by fewer "
        "samples (a frequency when < 1) are mapped directly; no_tree_map: return the graph without mutations.  "
        "Mutation IDs follow the row order.  Missing-data rows (allele '.', or missing=True with arrays) list the "
        "haplotypes with no call and come first among their site's rows; their node becomes the missingness node "
        "of the site's mutations.  hap_length (1..254, the reference's fixed 254 by default) sets the "
        "variants per haplotype window.  phased=False marks the graph unphased (carriers of an unphased individual fill "
        "its first haplotypes, as the reference's readers do, mut_iterator.cpp:404-410).");

    m.def("matmul", &matMul, py::arg("grg"), py::arg("input"), py::arg("direction"), py::arg("emit_all_nodes") = false,
          py::arg("by_individual") = false, py::arg("init") = py::none(), py::arg("miss") = py::none(),
          py::arg("order_preserving") = false,
          "pygrgl.matmul on the GPU: V(KxN)·X for UP, W(KxM)·X^T for DOWN (see the reference docstring, "
          "src/python/_grgl.cpp:1167-1229).  order_preserving=True sums every node in the reference's order, "
          "so float results match the CPU path bit for bit.");
    m.def("dot_product", &dotProduct, py::arg("grg"), py::arg("input"), py::arg("direction"),
          "DEPRECATED, use matmul instead.  1-D form of matmul.");

    m.def(
        "load_immutable_grg",
        [](const std::string& filename, bool loadUpEdges) {
            (void)loadUpEdges; // up edges are derived on demand
            auto g = std::make_shared<PyGRG>();
            {
                py::gil_scoped_release release;
                g->h = readGrgFile(filename);
            }
            return g;
        },
        py::arg("filename"), py::arg("load_up_edges") = true,
        "Load a .grg file (format 5.x, src/serialize.cpp:815-890 of the reference).");
    m.def(
        "save_grg",
        [](PyGRGPtr& g, const std::string& filename, bool allowSimplify) {
            py::gil_scoped_release release;
            writeGrgFile(g->H(), filename, allowSimplify);
        },
        py::arg("grg"), py::arg("filename"), py::arg("allow_simplify") = true,
        "Write the graph as a .grg file, format 5.3 (src/python/_grgl.cpp save_grg): with "
        "allow_simplify, nodes that only make the graph larger are removed as the reference's writer "
        "does (src/serialize.cpp:387-467); mutations of a graph built by map_mutations are written "
        "sorted by position and alleles.");

    m.def("grg_from_arrays", &fromArrays, py::arg("num_samples"), py::arg("ploidy"), py::arg("down_offsets"),
          py::arg("down_children"), py::arg("mutation_nodes"), py::arg("mutation_miss_nodes") = py::none(),
          py::arg("coal_counts") = py::none(), py::arg("positions") = py::none(),
          py::arg("ref_alleles") = py::none(), py::arg("alt_alleles") = py::none(),
          py::arg("populations") = py::none(), py::arg("population_ids") = py::none(),
          py::arg("individual_ids") = py::none(),
          "Build an immutable, ordered GRG from a down-edge CSR and the mutation->node map.");
    m.def(
        "grg_from_coalescent",
        [](uint32_t numSamples, uint32_t numSites, uint32_t numTrees, uint32_t ploidy, uint64_t seed, uint32_t threads) {
            SynthParams p;
            p.numSamples = numSamples;
            p.numSites = numSites;
            p.numTrees = numTrees;
            p.ploidy = ploidy;
            p.seed = seed;
            p.threads = threads;
            auto g = std::make_shared<PyGRG>();
            {
                py::gil_scoped_release release;
                g->h = synthCoalescentGrg(p);
            }
            return g;
        },
        py::arg("num_samples"), py::arg("num_sites"), py::arg("num_trees") = 1, py::arg("ploidy") = 2,
        py::arg("seed") = 20261203ULL, py::arg("threads") = 0,
        "Seeded synthetic GRG: Kingman coalescent trees (one per genome segment), mutations on branches with "
        "probability proportional to branch length, simplified as the reference writer does.");

    m.def(
        "_level_plan",
        [](PyGRGPtr& g, Direction dir) {
            const DirPlan p = dir == Direction::UP ? buildUpPlan(g->H()) : buildDownPlan(g->H());
            py::dict d;
            d["num_levels"] = p.numLevels;
            d["first_pulled_level"] = p.firstPulledLevel;
            d["level_start"] = toArray(p.levelStart);
            d["order"] = toArray(p.order);
            d["rank"] = toArray(p.rank);
            d["off"] = toArray(p.off);
            d["src"] = toArray(p.src);
            d["out_row"] = toArray(p.outRow);
            d["miss_row"] = toArray(p.missRow);
            d["coal_by_row"] = toArray(p.coalByRow);
            return d;
        },
        py::arg("grg"), py::arg("direction"),
        "The level-bucketed gather plan the device library builds for one direction (for inspection).");

    m.def(
        "_group_plans",
        [](PyGRGPtr& g, Direction dir, uint32_t heavyFanin, uint32_t ldsBytes, uint32_t maxEntries, uint32_t maxBlocks,
           uint32_t maxLevelRows) {
            const DirPlan p = dir == Direction::UP ? buildUpPlan(g->H()) : buildDownPlan(g->H());
            GroupLimits lim;
            lim.ldsBytes = ldsBytes;
            lim.maxEntries = maxEntries;
            lim.maxBlocks = maxBlocks;
            lim.maxLevelRows = maxLevelRows;
            const uint32_t regionEnd = groupRegionEnd(p, dir == Direction::DOWN);
            std::vector<GroupPlan> plans = buildGroupPlans(p, regionEnd, heavyFanin, lim);
            py::list out;
            for (const GroupPlan& gp : plans) {
                validateGroupPlan(p, gp, lim);
                py::dict d;
                d["slot_bytes"] = gp.slotBytes;
                d["heavy_fanin"] = gp.heavyFanin;
                d["level_begin"] = gp.levelBegin;
                d["level_end"] = gp.levelEnd;
                d["blocks"] = gp.blocks;
                d["lds_bytes"] = gp.ldsBytes;
                d["block_row"] = toArray(gp.blockRow);
                d["block_seg"] = toArray(gp.blockSeg);
                d["block_heavy"] = toArray(gp.blockHeavy);
                d["block_import"] = toArray(gp.blockImport);
                d["seg_row"] = toArray(gp.segRow);
                d["seg_heavy"] = toArray(gp.segHeavy);
                d["seg_heavy_ord"] = toArray(gp.segHeavyOrd);
                d["row"] = toArray(gp.row);
                d["off"] = toArray(gp.off);
                d["pre"] = toArray(gp.pre);
                d["src"] = toArray(gp.src);
                d["rem_off"] = toArray(gp.remOff);
                d["rem"] = toArray(gp.rem);
                d["heavy_row"] = toArray(gp.heavyRow);
                d["import_row"] = toArray(gp.importRow);
                out.append(d);
            }
            return out;
        },
        py::arg("grg"), py::arg("direction"), py::arg("heavy_fanin") = 64, py::arg("lds_bytes") = 65536,
        py::arg("max_entries") = 16384, py::arg("max_blocks") = 512, py::arg("max_level_rows") = 0xffffffffU,
        "The grouped-launch plans (one per row class: 4, 8, 16, 32, 64-byte slots) the device library "
        "builds for one direction, each validated (for inspection).");

    m.def(
        "_stack_offsets",
        [](const std::vector<uint32_t>& rows, size_t elemSize) {
            uint64_t total = 0;
            std::vector<uint32_t> off = stackOffsets(rows, elemSize, total);
            return std::make_pair(off, total);
        },
        py::arg("rows"), py::arg("elem_size"), "Stacked-row offsets the batch path assigns (for inspection).");

    m.def("set_device", [](int d) { g_device = d; }, py::arg("device"));
    m.def("get_device", []() { return g_device; });
    m.def("device_count", []() { return grgl_hip_device_count(); });

    m.attr("INVALID_NODE") = kInvalidNode;
    m.attr("COAL_COUNT_NOT_SET") = kCoalNotSet;
    m.attr("NO_UP_EDGES") = kNoUpEdges;
}
