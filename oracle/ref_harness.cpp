/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A C-callable harness around the reference's own CPU matmul.  oracle/Makefile compiles
 * this file together with the reference sources where they lie (src/grg.cpp,
 * src/mutation.cpp, third-party/libvbyte) into oracle/_ref/libgrgl_ref.so.  No reference
 * source is copied into this repository; the built library is git-ignored.
 *
 * The harness builds a CSRGRG exactly as the reference's loader leaves one
 * (include/grgl/grg.h:1232-1300: ordered nodes, samples first, sorted encoded down edges)
 * and calls GRG::matrixMultiplication<IO,NV,bit> (grg.h:1342-1517) on it, so tests can
 * compare the oracle and the GPU path with the reference bit for bit.
 */
#include "grgl/grg.h"
#include "grgl/mutation.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

using namespace grgl;

namespace {
thread_local std::string g_err;

struct RefGraph {
    std::shared_ptr<CSRGRG> grg;
};

template <typename T, typename NV, bool bit>
int runMatmul(RefGraph* h, const void* in, size_t inCols, size_t rows, int dir, void* out, size_t outSize,
              int emitAll, int byIndiv, const void* init, int initMode, void* miss) {
    h->grg->matrixMultiplication<T, NV, bit>((const T*)in, inCols, rows, (TraversalDirection)dir, (T*)out, outSize,
                                              emitAll != 0, byIndiv != 0, (const T*)init,
                                              (GRG::NodeInitEnum)initMode, (T*)miss);
    return 0;
}
} // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void* ref_graph_create(uint32_t numSamples, uint32_t numNodes, uint32_t ploidy, const uint64_t* downOff,
                       const uint32_t* downChild, uint32_t numMutations, const uint32_t* mutNode,
                       const uint32_t* mutMiss, const uint32_t* coal) {
    try {
        auto* h = new RefGraph();
        h->grg = std::make_shared<CSRGRG>(numSamples, numNodes, (uint16_t)ploidy, /*loadUpEdges=*/false);
        for (uint32_t n = 0; n < numNodes; n++) {
            if (downOff[n + 1] > downOff[n]) {
                std::vector<NodeID> kids(downChild + downOff[n], downChild + downOff[n + 1]);
                h->grg->m_downEdges.setData(n, kids);
            }
        }
        h->grg->m_downEdges.finalizeNodes();
        for (uint32_t m = 0; m < numMutations; m++) {
            const NodeID node = (mutNode[m] == 0x7fffffffU) ? INVALID_NODE_ID : mutNode[m];
            const NodeID miss = (mutMiss == nullptr || mutMiss[m] == 0x7fffffffU) ? INVALID_NODE_ID : mutMiss[m];
            h->grg->addMutation(Mutation((BpPosition)m, "G", "A"), node, miss);
        }
        if (coal != nullptr) {
            for (uint32_t n = numSamples; n < numNodes; n++) {
                h->grg->setNumIndividualCoalsGrow(n, coal[n]);
            }
        }
        return h;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void ref_graph_destroy(void* handle) { delete (RefGraph*)handle; }

/* dtype codes as in oracle/grg_oracle.h.  Returns 0, 1 (ApiMisuseFailure) or 2 (other). */
int ref_matmul(void* handle, int dtype, const void* in, size_t inCols, size_t rows, int dir, void* out,
               size_t outSize, int emitAll, int byIndiv, const void* init, int initMode, void* miss) {
    auto* h = (RefGraph*)handle;
    try {
        switch (dtype) {
        case 0: return runMatmul<double, double, false>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        case 1: return runMatmul<float, float, false>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        case 2: return runMatmul<int64_t, int64_t, false>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        case 3: return runMatmul<int32_t, int32_t, false>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        case 4: return runMatmul<int16_t, int16_t, false>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        case 5: return runMatmul<int8_t, int8_t, false>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        case 6: return runMatmul<uint64_t, uint64_t, false>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        case 7: return runMatmul<uint32_t, uint32_t, false>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        case 8: return runMatmul<uint16_t, uint16_t, false>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        case 9: return runMatmul<uint8_t, uint8_t, false>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        case 10: return runMatmul<bool, uint8_t, true>(h, in, inCols, rows, dir, out, outSize, emitAll, byIndiv, init, initMode, miss);
        default: g_err = "unsupported dtype"; return 1;
        }
    } catch (const ApiMisuseFailure& e) {
        g_err = e.what();
        return 1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 2;
    }
}

uint32_t ref_num_edges(void* handle) { return (uint32_t)((RefGraph*)handle)->grg->numEdges(); }

/*
 * The edge section of a .grg file as the reference's own CSR storage produces it
 * (EagerCSREdges32::setData + flushBuckets/flushIndexes, csr_storage.h:440-553,
 * src/serialize.cpp:450,575): encoded child lists, then numNodes + 1 u64 byte offsets.
 * Returns the number of edge bytes; offsets go to outIndex.  outBytes may be NULL to
 * query the size.
 */
/*
 * The 24 bytes the reference writes for one mutation record of a .grg file
 * (writeScalar<Mutation>, src/serialize.cpp:618; layout include/grgl/mutation.h:248-261):
 * the raw object of Mutation(position, alt, ref, time).  Returns 1 when the alleles are
 * stored inline (isShort), 0 when the first 8 bytes hold a heap pointer with bit 0 set and
 * the allele storage string follows the records (returned in outStorage, capacity bytes).
 */
int ref_mutation_record(uint64_t position, const char* ref, const char* alt, uint32_t time, uint8_t* out24,
                        char* outStorage, size_t capacity) {
    Mutation mut((BpPosition)position, std::string(alt), std::string(ref), time);
    std::memcpy(out24, (const void*)&mut, sizeof(Mutation));
    if (mut.isShort()) return 1;
    const std::string st = mut.alleleStorageString();
    if (outStorage != nullptr && st.size() < capacity) std::memcpy(outStorage, st.c_str(), st.size() + 1);
    return 0;
}

size_t ref_encode_edges(uint32_t numNodes, const uint64_t* downOff, const uint32_t* downChild, uint8_t* outBytes,
                        size_t capacity, uint64_t* outIndex) {
    EagerCSREdges32 csr(numNodes);
    for (uint32_t n = 0; n < numNodes; n++) {
        if (downOff[n + 1] > downOff[n]) {
            std::vector<NodeID> kids(downChild + downOff[n], downChild + downOff[n + 1]);
            csr.setData(n, kids);
        }
    }
    csr.finalizeNodes();
    std::ostringstream bucket, index;
    csr.flushBuckets(bucket, /*clearData=*/false);
    csr.flushIndexes(index, /*clearData=*/false);
    const std::string b = bucket.str(), i = index.str();
    if (outBytes != nullptr && b.size() <= capacity) std::memcpy(outBytes, b.data(), b.size());
    if (outIndex != nullptr) std::memcpy(outIndex, i.data(), i.size());
    return b.size();
}
}
