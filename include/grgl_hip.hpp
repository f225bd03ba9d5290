/*
 * grgl_hip.hpp — C++ face of the C ABI in grgl_hip.h, with the reference's C++ matmul API.
 *
 * grgl_hip::DeviceGRG mirrors, on a device-resident graph:
 *   GRG::matMul<T>(inputMatrix, numRows, direction)            include/grgl/grg.h:705-716
 *   GRG::matrixMultiplication<IOType,NodeValueType,useBitVector>(...)  grg.h:656-667
 * with the same arguments, the same `+=` into a caller-zeroed output, and the same
 * exception type for misuse (ApiMisuseFailure, a std::runtime_error; common.h:45-49).
 * Buffers are host memory (the call stages them and synchronises).  Header-only; link
 * against libgrgl_hip.so.
 */
#ifndef GRGL_HIP_HPP
#define GRGL_HIP_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "grgl_hip.h"

namespace grgl_hip {

enum TraversalDirection { DIRECTION_DOWN = GRGL_DIRECTION_DOWN, DIRECTION_UP = GRGL_DIRECTION_UP };

enum NodeInitEnum {
    NIE_ZERO = GRGL_INIT_ZERO,
    NIE_VECTOR = GRGL_INIT_VECTOR,
    NIE_MATRIX = GRGL_INIT_MATRIX,
    NIE_XTX = GRGL_INIT_XTX,
};

class ApiMisuseFailure : public std::runtime_error {
public:
    explicit ApiMisuseFailure(const std::string& m) : std::runtime_error(m) {}
};

class DeviceFailure : public std::runtime_error {
public:
    explicit DeviceFailure(const std::string& m) : std::runtime_error(m) {}
};

inline void check(grgl_status st) {
    if (st == GRGL_OK) return;
    const std::string msg = grgl_hip_last_error();
    if (st == GRGL_ERR_MISUSE) throw ApiMisuseFailure(msg);
    throw DeviceFailure(msg);
}

template <typename T> struct DTypeOf;
template <> struct DTypeOf<double> { static constexpr grgl_dtype value = GRGL_F64; };
template <> struct DTypeOf<float> { static constexpr grgl_dtype value = GRGL_F32; };
template <> struct DTypeOf<int64_t> { static constexpr grgl_dtype value = GRGL_I64; };
template <> struct DTypeOf<int32_t> { static constexpr grgl_dtype value = GRGL_I32; };
template <> struct DTypeOf<int16_t> { static constexpr grgl_dtype value = GRGL_I16; };
template <> struct DTypeOf<int8_t> { static constexpr grgl_dtype value = GRGL_I8; };
template <> struct DTypeOf<uint64_t> { static constexpr grgl_dtype value = GRGL_U64; };
template <> struct DTypeOf<uint32_t> { static constexpr grgl_dtype value = GRGL_U32; };
template <> struct DTypeOf<uint16_t> { static constexpr grgl_dtype value = GRGL_U16; };
template <> struct DTypeOf<uint8_t> { static constexpr grgl_dtype value = GRGL_U8; };
template <> struct DTypeOf<bool> { static constexpr grgl_dtype value = GRGL_BOOL; };

/* An immutable, ordered GRG resident on one device (the content of a CSRGRG). */
class DeviceGRG {
public:
    DeviceGRG(uint32_t numSamples, uint32_t ploidy, const std::vector<uint64_t>& downOffsets,
              const std::vector<uint32_t>& downChildren, const std::vector<uint32_t>& mutationNodes,
              const std::vector<uint32_t>& mutationMissNodes = {}, const std::vector<uint32_t>& coalCounts = {},
              int device = 0)
        : m_numSamples(numSamples),
          m_ploidy(ploidy),
          m_numNodes(downOffsets.empty() ? 0 : (uint32_t)(downOffsets.size() - 1)),
          m_numMutations((uint32_t)mutationNodes.size()) {
        if (downOffsets.empty()) throw ApiMisuseFailure("downOffsets must have numNodes + 1 entries");
        grgl_graph_desc d{};
        d.num_samples = numSamples;
        d.num_nodes = m_numNodes;
        d.ploidy = ploidy;
        d.num_mutations = m_numMutations;
        d.down_offsets = downOffsets.data();
        d.down_children = downChildren.data();
        d.mutation_nodes = mutationNodes.data();
        d.mutation_miss_nodes = mutationMissNodes.empty() ? nullptr : mutationMissNodes.data();
        d.coal_counts = coalCounts.empty() ? nullptr : coalCounts.data();
        check(grgl_hip_graph_create(&d, device, &m_graph));
    }
    DeviceGRG(const DeviceGRG&) = delete;
    DeviceGRG& operator=(const DeviceGRG&) = delete;
    ~DeviceGRG() { grgl_hip_graph_destroy(m_graph); }

    /* loadGPUGRGFromDisk (include/grgl/cuda/gpu_grg.h:482-489): a graph from a levelized-graph
     * file written by store() or grgl_hip_levelize_to_file. */
    static std::unique_ptr<DeviceGRG> load(const std::string& path, int device = 0) {
        grgl_graph* g = nullptr;
        check(grgl_hip_graph_load(path.c_str(), device, &g));
        grgl_graph_info info{};
        const grgl_status st = grgl_hip_graph_info(g, &info);
        if (st != GRGL_OK) {
            grgl_hip_graph_destroy(g);
            check(st);
        }
        return std::unique_ptr<DeviceGRG>(new DeviceGRG(g, info));
    }
    /* storeGPUGRGToDisk (include/grgl/cuda/gpu_grg.h:473-480). */
    void store(const std::string& path) const { check(grgl_hip_graph_store(m_graph, path.c_str())); }

    uint32_t numSamples() const { return m_numSamples; }
    uint32_t numIndividuals() const { return m_numSamples / m_ploidy; }
    uint32_t numNodes() const { return m_numNodes; }
    uint32_t numMutations() const { return m_numMutations; }
    uint32_t getPloidy() const { return m_ploidy; }
    grgl_graph* handle() const { return m_graph; }

    /* GRG::matrixMultiplication (grg.h:656-667): output is added to. */
    template <typename IOType, typename NodeValueType = IOType, bool useBitVector = false>
    void matrixMultiplication(const IOType* inputMatrix, size_t inputCols, size_t inputRows,
                              TraversalDirection direction, IOType* outputMatrix, size_t outputSize,
                              bool emitAllNodes = false, bool byIndividual = false,
                              const IOType* initMatrix = nullptr, NodeInitEnum nodeInit = NIE_ZERO,
                              IOType* missMatrix = nullptr) {
        static_assert(std::is_same<IOType, NodeValueType>::value || useBitVector,
                      "NodeValueType differs from IOType only for the bit-vector (bool) form");
        check(grgl_hip_matmul(m_graph, DTypeOf<IOType>::value, inputMatrix, inputCols, inputRows,
                              (grgl_direction)direction, outputMatrix, outputSize, emitAllNodes ? 1 : 0,
                              byIndividual ? 1 : 0, initMatrix, (grgl_init_mode)nodeInit, missMatrix,
                              GRGL_FLAG_HOST_PTRS | GRGL_FLAG_ACCUMULATE, nullptr));
    }

    /* GRG::matMul<T> (grg.h:705-716). */
    template <typename T>
    std::vector<T> matMul(const std::vector<T>& inputMatrix, const size_t numRows, TraversalDirection direction) {
        if (numRows == 0 || (inputMatrix.size() % numRows != 0)) {
            throw ApiMisuseFailure("inputMatrix must be divisible by numRows");
        }
        const size_t numCols = inputMatrix.size() / numRows;
        const size_t outSize = numRows * (direction == DIRECTION_DOWN ? numSamples() : numMutations());
        std::vector<T> result(outSize);
        matrixMultiplication<T, T, false>(inputMatrix.data(), numCols, numRows, direction, result.data(), outSize);
        return result;
    }

private:
    DeviceGRG(grgl_graph* g, const grgl_graph_info& info)
        : m_graph(g),
          m_numSamples(info.num_samples),
          m_ploidy(info.ploidy),
          m_numNodes(info.num_nodes),
          m_numMutations(info.num_mutations) {}

    grgl_graph* m_graph = nullptr;
    uint32_t m_numSamples, m_ploidy, m_numNodes, m_numMutations;
};

} // namespace grgl_hip

#endif /* GRGL_HIP_HPP */
