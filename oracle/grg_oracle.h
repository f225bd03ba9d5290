/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of GRGL's CPU graph matrix multiplication
 * (include/grgl/grg.h:1342-1517, include/grgl/internal_mult.h:42-307), used by the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark as the
 * checker.  The product path (grgl_amd, libgrgl_hip.so) never links, loads or calls
 * anything in this directory.
 *
 * Parity pinning (each function below is checked by at least one of):
 *   - the reference's own golden vectors (test/unit/test_grg.cpp:41-124, :258-292,
 *     :333-367; test/endtoend/expect/test-200-samples.freq.txt);
 *   - bit-exact agreement with the reference's own matrixMultiplication compiled from
 *     its sources by oracle/Makefile into oracle/_ref (all dtypes, both directions,
 *     init/miss/by_individual/emit_all_nodes);
 *   - closed forms: UP(ones) = subtree sample counts, dense brute force V @ X,
 *     and the adjoint identity sum(UP(V) * W) == sum(V * DOWN(W)) in int64.
 */
#ifndef GRG_ORACLE_H
#define GRG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/grgl/grgnode.h:57-58 */
#define ORACLE_INVALID_NODE 0x7fffffffU
/* include/grgl/node_data.h:31 */
#define ORACLE_COAL_NOT_SET 0xffffffffU

/* include/grgl/visitor.h:26 */
enum { ORACLE_DOWN = 1, ORACLE_UP = 2 };

/* include/grgl/grg.h:646-651 */
enum { ORACLE_INIT_ZERO = 0, ORACLE_INIT_VECTOR = 1, ORACLE_INIT_MATRIX = 2, ORACLE_INIT_XTX = 3 };

/* dtype codes shared with the product's C ABI (include/grgl_hip.h) */
enum {
    ORACLE_F64 = 0, ORACLE_F32 = 1, ORACLE_I64 = 2, ORACLE_I32 = 3, ORACLE_I16 = 4, ORACLE_I8 = 5,
    ORACLE_U64 = 6, ORACLE_U32 = 7, ORACLE_U16 = 8, ORACLE_U8 = 9, ORACLE_BOOL = 10
};

/* Status codes: 0 ok, 1 API misuse (the reference's ApiMisuseFailure), 2 internal. */
enum { ORACLE_OK = 0, ORACLE_ERR_MISUSE = 1, ORACLE_ERR_INTERNAL = 2 };

/*
 * An immutable, ordered GRG (the reference's CSRGRG, include/grgl/grg.h:1232-1300):
 * nodes 0..num_samples-1 are the samples, every child ID is smaller than its parent's,
 * each node's children are sorted ascending (EagerCSREdges32 is Sorted, csr_storage.h:580).
 */
typedef struct {
    uint32_t num_samples;
    uint32_t num_nodes;
    uint32_t ploidy;
    uint32_t num_mutations;
    const uint64_t* down_off;   /* num_nodes + 1 */
    const uint32_t* down_child; /* down_off[num_nodes] */
    const uint32_t* mut_node;   /* per MutationId: node or ORACLE_INVALID_NODE */
    const uint32_t* mut_miss;   /* per MutationId: missingness node or invalid; may be NULL */
    const uint32_t* coal;       /* per node: individual coalescence count; may be NULL */
} oracle_graph;

/*
 * GRG::matrixMultiplication (grg.h:1342-1517), with the caller-owned output accumulated
 * into (+=), exactly as the reference does.  `dtype` selects IOType == NodeValueType, or
 * the bool/bit-vector GF(2) variant for ORACLE_BOOL (internal_mult.h:111-128).
 * Returns ORACLE_OK or ORACLE_ERR_MISUSE (message via oracle_last_error()).
 */
int oracle_matmul(const oracle_graph* g, int dtype, const void* input, size_t input_cols, size_t input_rows,
                  int direction, void* output, size_t output_size, int emit_all_nodes, int by_individual,
                  const void* init, int init_mode, void* miss);

const char* oracle_last_error(void);

/*
 * "Faithful" timing mode of the same computation: every node's child list is stored
 * delta-varint (LEB128) encoded and decoded into a freshly allocated buffer on every
 * visit, as CSRStorageImm::getData does (csr_storage.h:407-429, vbyte.cc:422-446).
 * Only f64/f32/i64/i32 are provided; used for CPU-baseline timing when the compiled
 * reference is not available.
 */
typedef struct oracle_encoded_graph oracle_encoded_graph;
oracle_encoded_graph* oracle_encode(const oracle_graph* g);
void oracle_encoded_free(oracle_encoded_graph* eg);
int oracle_matmul_faithful(const oracle_encoded_graph* eg, int dtype, const void* input, size_t input_rows,
                           int direction, void* output);

#ifdef __cplusplus
}
#endif
#endif
