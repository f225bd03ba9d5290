/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.  See grg_oracle.h for what this is, how it is
 * pinned, and who may call it.  Never linked into the product library.
 *
 * Every loop below follows GRG::matrixMultiplication (include/grgl/grg.h:1342-1517) in
 * the same phase order and the same per-node summation order, so float results are
 * bit-identical to the reference CPU path, not merely close:
 *   - node values are node-major with the K rows contiguous (grg.h:1362-1367);
 *   - DOWN seeds in (node, mutId) order with unmapped mutations first (grg.h:1406-1427,
 *     NodeMutMissLt grg.h:855-865), then pushes parent -> child for parents in
 *     descending ID order and children ascending (grg.h:1428-1431, :1312-1322);
 *   - UP seeds samples (grg.h:1456-1469), then pulls children ascending for nodes
 *     numSamples..numNodes-1 (grg.h:1470-1474, :1324-1334);
 *   - read-outs accumulate into the caller's output (grg.h:1440-1453, :1483-1516).
 * Integer adds wrap modulo 2^bits (done in unsigned arithmetic so C has no UB), which is
 * what the reference's `+=` on fixed-width integers compiles to (internal_mult.h:42-49).
 */
#include "grg_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static _Thread_local char g_err[512];

const char* oracle_last_error(void) { return g_err; }

static int fail(int code, const char* msg) {
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return code;
}

/* validateMatMulInputs (internal_mult.h:130-158). */
static int validate(const oracle_graph* g, size_t input_cols, size_t input_rows, int direction,
                    size_t output_size, int emit_all, int by_indiv, size_t output_cols) {
    if (output_size % input_rows != 0) {
        return fail(ORACLE_ERR_INTERNAL, "outputSize not divisible by inputRows");
    }
    const size_t expect_sample_cols = by_indiv ? (g->num_samples / g->ploidy) : g->num_samples;
    if (direction == ORACLE_DOWN) {
        if (input_cols != g->num_mutations) {
            return fail(ORACLE_ERR_MISUSE, "Input vector is not size M (numMutations)");
        }
        if (!emit_all && output_cols != expect_sample_cols) {
            return fail(ORACLE_ERR_MISUSE,
                        "Output vector is not size N (numSamples, or numIndividuals depending on flags)");
        }
    } else {
        if (input_cols != expect_sample_cols) {
            return fail(ORACLE_ERR_MISUSE,
                        "Input vector is not size N (numSamples, or numIndividuals depending on flags)");
        }
        if (!emit_all && output_cols != g->num_mutations) {
            return fail(ORACLE_ERR_MISUSE, "Output vector is not size M (numMutations)");
        }
    }
    if (emit_all && output_cols != g->num_nodes) {
        return fail(ORACLE_ERR_MISUSE, "Output vector is not size numNodes");
    }
    return ORACLE_OK;
}

/*
 * The mutation map in the order getNodesAndMutations() yields it: sorted by
 * (node, mutId) with INVALID_NODE_ID entries first (grg.cpp:265-301, grg.h:855-865).
 * A stable bucket sort by node gives the same order because mutIds are unique.
 */
static uint32_t* sorted_mutation_order(const oracle_graph* g) {
    const uint32_t M = g->num_mutations, V = g->num_nodes;
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (M ? M : 1));
    size_t* start = (size_t*)calloc((size_t)V + 2, sizeof(size_t));
    if (!order || !start) {
        free(order);
        free(start);
        return NULL;
    }
    for (uint32_t m = 0; m < M; m++) {
        const uint32_t n = g->mut_node[m];
        start[(n == ORACLE_INVALID_NODE) ? 0 : (size_t)n + 1]++;
    }
    size_t acc = 0;
    for (size_t b = 0; b < (size_t)V + 1; b++) {
        const size_t c = start[b];
        start[b] = acc;
        acc += c;
    }
    for (uint32_t m = 0; m < M; m++) {
        const uint32_t n = g->mut_node[m];
        order[start[(n == ORACLE_INVALID_NODE) ? 0 : (size_t)n + 1]++] = m;
    }
    free(start);
    return order;
}

/* Element add of internal_mult.h:42-49 per dtype. */
#define ADD_FLOAT(a, b) ((a) + (b))
#define ADD_WRAP(UT, T, a, b) ((T)(UT)((UT)(a) + (UT)(b)))

#define DEFINE_ORACLE_MATMUL(SUF, T, ADDEXPR)                                                                    \
    static inline T add_##SUF(T a, T b) { return ADDEXPR; }                                                     \
    static int matmul_##SUF(const oracle_graph* g, const T* in, size_t in_cols, size_t K, int dir, T* out,      \
                            size_t out_size, int emit_all, int by_indiv, const T* init, int init_mode, T* miss) { \
        const size_t out_cols = out_size / K;                                                                  \
        const size_t V = g->num_nodes, N = g->num_samples, P = g->ploidy;                                      \
        T* val = (T*)calloc(V * K ? V * K : 1, sizeof(T));                                                     \
        if (!val) return fail(ORACLE_ERR_INTERNAL, "out of memory");                                           \
        /* node initialisation, grg.h:1369-1402 */                                                             \
        if (init_mode == ORACLE_INIT_XTX) {                                                                    \
            for (size_t i = 0; i < V; i++) {                                                                   \
                uint32_t coal = 0;                                                                             \
                if (i >= N) coal = g->coal ? g->coal[i] : ORACLE_COAL_NOT_SET;                                 \
                if (coal == ORACLE_COAL_NOT_SET) {                                                             \
                    free(val);                                                                                 \
                    return fail(ORACLE_ERR_MISUSE,                                                             \
                                "Coalescent counts not available for this GRG (init=XTX won't work)");         \
                }                                                                                              \
                const T c2 = (T)(coal * 2U);                                                                   \
                for (size_t r = 0; r < K; r++) val[i * K + r] = add_##SUF(val[i * K + r], c2);                 \
            }                                                                                                  \
        } else if (init_mode == ORACLE_INIT_VECTOR) {                                                          \
            for (size_t i = 0; i < V; i++)                                                                     \
                for (size_t r = 0; r < K; r++) val[i * K + r] = add_##SUF(val[i * K + r], init[r]);            \
        } else if (init_mode == ORACLE_INIT_MATRIX) {                                                          \
            for (size_t i = 0; i < V; i++)                                                                     \
                for (size_t r = 0; r < K; r++)                                                                 \
                    val[i * K + r] = add_##SUF(val[i * K + r], init[r * V + i]);                               \
        }                                                                                                      \
        uint32_t* order = sorted_mutation_order(g);                                                            \
        if (!order) {                                                                                          \
            free(val);                                                                                         \
            return fail(ORACLE_ERR_INTERNAL, "out of memory");                                                 \
        }                                                                                                      \
        const uint32_t M = g->num_mutations;                                                                   \
        if (dir == ORACLE_DOWN) {                                                                              \
            /* seeds, grg.h:1406-1427 */                                                                       \
            for (uint32_t t = 0; t < M; t++) {                                                                 \
                const uint32_t m = order[t];                                                                   \
                const uint32_t node = g->mut_node[m];                                                          \
                const uint32_t mn = g->mut_miss ? g->mut_miss[m] : ORACLE_INVALID_NODE;                        \
                if (node != ORACLE_INVALID_NODE)                                                               \
                    for (size_t r = 0; r < K; r++)                                                             \
                        val[node * K + r] = add_##SUF(val[node * K + r], in[r * in_cols + m]);                 \
                if (miss && mn != ORACLE_INVALID_NODE)                                                         \
                    for (size_t r = 0; r < K; r++)                                                             \
                        val[(size_t)mn * K + r] = add_##SUF(val[(size_t)mn * K + r], miss[r * in_cols + m]);   \
            }                                                                                                  \
            /* push down, parents in descending ID order, grg.h:1428-1431, :1312-1322 */                       \
            for (size_t i = V; i > 0; i--) {                                                                   \
                const size_t p = i - 1;                                                                        \
                for (uint64_t e = g->down_off[p]; e < g->down_off[p + 1]; e++) {                               \
                    const size_t c = g->down_child[e];                                                         \
                    for (size_t r = 0; r < K; r++) val[c * K + r] = add_##SUF(val[c * K + r], val[p * K + r]); \
                }                                                                                              \
            }                                                                                                  \
            if (!emit_all) { /* read-out, grg.h:1440-1453 */                                                   \
                for (size_t r = 0; r < K; r++)                                                                 \
                    for (size_t s = 0; s < N; s++) {                                                           \
                        const size_t col = by_indiv ? s / P : s;                                               \
                        out[r * out_cols + col] = add_##SUF(out[r * out_cols + col], val[s * K + r]);          \
                    }                                                                                          \
            }                                                                                                  \
        } else {                                                                                               \
            /* seed samples, grg.h:1456-1469 */                                                                \
            for (size_t s = 0; s < N; s++) {                                                                   \
                const size_t col = by_indiv ? s / P : s;                                                       \
                for (size_t r = 0; r < K; r++) val[s * K + r] = add_##SUF(val[s * K + r], in[r * in_cols + col]); \
            }                                                                                                  \
            /* pull up, children ascending, nodes numSamples.. (samplesAreOrdered), grg.h:1470-1474 */         \
            for (size_t n = N; n < V; n++) {                                                                   \
                for (uint64_t e = g->down_off[n]; e < g->down_off[n + 1]; e++) {                               \
                    const size_t c = g->down_child[e];                                                         \
                    for (size_t r = 0; r < K; r++) val[n * K + r] = add_##SUF(val[n * K + r], val[c * K + r]); \
                }                                                                                              \
            }                                                                                                  \
            if (!emit_all) { /* read-out, grg.h:1483-1505 */                                                   \
                for (size_t r = 0; r < K; r++)                                                                 \
                    for (uint32_t t = 0; t < M; t++) {                                                         \
                        const uint32_t m = order[t];                                                           \
                        const uint32_t node = g->mut_node[m];                                                  \
                        const uint32_t mn = g->mut_miss ? g->mut_miss[m] : ORACLE_INVALID_NODE;                \
                        if (node != ORACLE_INVALID_NODE)                                                       \
                            out[r * out_cols + m] = add_##SUF(out[r * out_cols + m], val[(size_t)node * K + r]); \
                        if (miss && mn != ORACLE_INVALID_NODE)                                                 \
                            miss[r * out_cols + m] = add_##SUF(miss[r * out_cols + m], val[(size_t)mn * K + r]); \
                    }                                                                                          \
            }                                                                                                  \
        }                                                                                                      \
        if (emit_all) { /* grg.h:1507-1516 */                                                                  \
            for (size_t r = 0; r < K; r++)                                                                     \
                for (size_t n = 0; n < V; n++)                                                                 \
                    out[r * out_cols + n] = add_##SUF(out[r * out_cols + n], val[n * K + r]);                  \
        }                                                                                                      \
        free(order);                                                                                           \
        free(val);                                                                                             \
        return ORACLE_OK;                                                                                      \
    }

DEFINE_ORACLE_MATMUL(f64, double, ADD_FLOAT(a, b))
DEFINE_ORACLE_MATMUL(f32, float, ADD_FLOAT(a, b))
DEFINE_ORACLE_MATMUL(i64, int64_t, ADD_WRAP(uint64_t, int64_t, a, b))
DEFINE_ORACLE_MATMUL(i32, int32_t, ADD_WRAP(uint32_t, int32_t, a, b))
DEFINE_ORACLE_MATMUL(i16, int16_t, ADD_WRAP(uint16_t, int16_t, a, b))
DEFINE_ORACLE_MATMUL(i8, int8_t, ADD_WRAP(uint8_t, int8_t, a, b))
DEFINE_ORACLE_MATMUL(u64, uint64_t, ADD_WRAP(uint64_t, uint64_t, a, b))
DEFINE_ORACLE_MATMUL(u32, uint32_t, ADD_WRAP(uint32_t, uint32_t, a, b))
DEFINE_ORACLE_MATMUL(u16, uint16_t, ADD_WRAP(uint16_t, uint16_t, a, b))
DEFINE_ORACLE_MATMUL(u8, uint8_t, ADD_WRAP(uint8_t, uint8_t, a, b))
/* bool: the bit-vector GF(2) path (internal_mult.h:111-128, :207-241) holds one bit per
 * (node,row); a byte per bit with XOR is the same arithmetic.  Seeding flips the bit when
 * the input is true; read-out sets out = (out != bit). Both are XOR on 0/1 bytes. */
DEFINE_ORACLE_MATMUL(b8, uint8_t, (uint8_t)(((a) ^ (b)) & 1U))

int oracle_matmul(const oracle_graph* g, int dtype, const void* input, size_t input_cols, size_t input_rows,
                  int direction, void* output, size_t output_size, int emit_all_nodes, int by_individual,
                  const void* init, int init_mode, void* miss) {
    g_err[0] = 0;
    /* release_asserts of grg.h:1353-1354 */
    if (input_cols == 0 || input_rows == 0) {
        return fail(ORACLE_ERR_INTERNAL, "inputCols and inputRows must be > 0");
    }
    const size_t output_cols = output_size / input_rows;
    int rc = validate(g, input_cols, input_rows, direction, output_size, emit_all_nodes, by_individual, output_cols);
    if (rc != ORACLE_OK) {
        return rc;
    }
    if (dtype == ORACLE_BOOL && init_mode != ORACLE_INIT_ZERO) {
        return fail(ORACLE_ERR_MISUSE, "Node initialization and bit vector type cannot be mixed");
    }
    switch (dtype) {
#define CASE(CODE, SUF, T)                                                                                       \
    case CODE:                                                                                                  \
        return matmul_##SUF(g, (const T*)input, input_cols, input_rows, direction, (T*)output, output_size,    \
                            emit_all_nodes, by_individual, (const T*)init, init_mode, (T*)miss);
        CASE(ORACLE_F64, f64, double)
        CASE(ORACLE_F32, f32, float)
        CASE(ORACLE_I64, i64, int64_t)
        CASE(ORACLE_I32, i32, int32_t)
        CASE(ORACLE_I16, i16, int16_t)
        CASE(ORACLE_I8, i8, int8_t)
        CASE(ORACLE_U64, u64, uint64_t)
        CASE(ORACLE_U32, u32, uint32_t)
        CASE(ORACLE_U16, u16, uint16_t)
        CASE(ORACLE_U8, u8, uint8_t)
        CASE(ORACLE_BOOL, b8, uint8_t)
#undef CASE
    default: return fail(ORACLE_ERR_MISUSE, "Unsupported dtype");
    }
}

/* ---------------------------------------------------------------------------------------
 * Faithful timing mode: per-node delta-varint lists decoded into a fresh allocation on
 * every visit (csr_storage.h:407-429; LEB128 codec of vbyte.cc:149-192 / :422-446:
 * 7 payload bits per byte, high bit set on all but the last byte, sorted lists store
 * differences from the previous value with the first value relative to 0).
 * ------------------------------------------------------------------------------------- */
struct oracle_encoded_graph {
    oracle_graph g;
    uint64_t* node_byte_off; /* num_nodes + 1 */
    uint8_t* bytes;
};

static size_t put_varint(uint8_t* out, uint32_t v) {
    size_t n = 0;
    while (v >= 0x80U) {
        out[n++] = (uint8_t)(v | 0x80U);
        v >>= 7;
    }
    out[n++] = (uint8_t)v;
    return n;
}

static size_t get_varint(const uint8_t* in, uint32_t* v) {
    uint32_t r = 0;
    unsigned shift = 0;
    size_t n = 0;
    for (;;) {
        const uint8_t b = in[n++];
        r |= (uint32_t)(b & 0x7fU) << shift;
        if (!(b & 0x80U)) break;
        shift += 7;
    }
    *v = r;
    return n;
}

oracle_encoded_graph* oracle_encode(const oracle_graph* g) {
    oracle_encoded_graph* eg = (oracle_encoded_graph*)calloc(1, sizeof(*eg));
    if (!eg) return NULL;
    eg->g = *g;
    const uint64_t E = g->down_off[g->num_nodes];
    eg->node_byte_off = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)g->num_nodes + 1));
    eg->bytes = (uint8_t*)malloc((size_t)(E + g->num_nodes) * 5 + 1);
    if (!eg->node_byte_off || !eg->bytes) {
        oracle_encoded_free(eg);
        return NULL;
    }
    uint64_t pos = 0;
    for (uint32_t n = 0; n < g->num_nodes; n++) {
        eg->node_byte_off[n] = pos;
        const uint64_t b = g->down_off[n], e = g->down_off[n + 1];
        if (e > b) {
            pos += put_varint(eg->bytes + pos, (uint32_t)(e - b));
            uint32_t prev = 0;
            for (uint64_t i = b; i < e; i++) {
                pos += put_varint(eg->bytes + pos, g->down_child[i] - prev);
                prev = g->down_child[i];
            }
        }
    }
    eg->node_byte_off[g->num_nodes] = pos;
    return eg;
}

void oracle_encoded_free(oracle_encoded_graph* eg) {
    if (!eg) return;
    free(eg->node_byte_off);
    free(eg->bytes);
    free(eg);
}

static uint32_t* decode_children(const oracle_encoded_graph* eg, uint32_t n, uint32_t* count) {
    const uint64_t b = eg->node_byte_off[n], e = eg->node_byte_off[n + 1];
    *count = 0;
    if (e <= b) return NULL;
    const uint8_t* p = eg->bytes + b;
    uint32_t cnt;
    p += get_varint(p, &cnt);
    uint32_t* out = (uint32_t*)malloc(sizeof(uint32_t) * cnt);
    uint32_t prev = 0;
    for (uint32_t i = 0; i < cnt; i++) {
        uint32_t d;
        p += get_varint(p, &d);
        prev += d;
        out[i] = prev;
    }
    *count = cnt;
    return out;
}

#define DEFINE_FAITHFUL(SUF, T)                                                                                  \
    static int faithful_##SUF(const oracle_encoded_graph* eg, const T* in, size_t K, int dir, T* out) {        \
        const oracle_graph* g = &eg->g;                                                                        \
        const size_t V = g->num_nodes, N = g->num_samples;                                                     \
        const uint32_t M = g->num_mutations;                                                                   \
        T* val = (T*)calloc(V * K, sizeof(T));                                                                 \
        uint32_t* order = sorted_mutation_order(g);                                                            \
        if (!val || !order) {                                                                                  \
            free(val);                                                                                         \
            free(order);                                                                                       \
            return fail(ORACLE_ERR_INTERNAL, "out of memory");                                                 \
        }                                                                                                      \
        if (dir == ORACLE_DOWN) {                                                                              \
            for (uint32_t t = 0; t < M; t++) {                                                                 \
                const uint32_t m = order[t], node = g->mut_node[m];                                            \
                if (node != ORACLE_INVALID_NODE)                                                               \
                    for (size_t r = 0; r < K; r++) val[(size_t)node * K + r] += in[r * M + m];                 \
            }                                                                                                  \
            for (size_t i = V; i > 0; i--) {                                                                   \
                uint32_t cnt;                                                                                  \
                uint32_t* ch = decode_children(eg, (uint32_t)(i - 1), &cnt);                                   \
                const T* src = val + (i - 1) * K;                                                              \
                for (uint32_t j = 0; j < cnt; j++) {                                                           \
                    T* dst = val + (size_t)ch[j] * K;                                                          \
                    for (size_t r = 0; r < K; r++) dst[r] += src[r];                                           \
                }                                                                                              \
                free(ch);                                                                                      \
            }                                                                                                  \
            for (size_t r = 0; r < K; r++)                                                                     \
                for (size_t s = 0; s < N; s++) out[r * N + s] += val[s * K + r];                               \
        } else {                                                                                               \
            for (size_t s = 0; s < N; s++)                                                                     \
                for (size_t r = 0; r < K; r++) val[s * K + r] += in[r * N + s];                                \
            for (size_t n = N; n < V; n++) {                                                                   \
                uint32_t cnt;                                                                                  \
                uint32_t* ch = decode_children(eg, (uint32_t)n, &cnt);                                         \
                T* dst = val + n * K;                                                                          \
                for (uint32_t j = 0; j < cnt; j++) {                                                           \
                    const T* src = val + (size_t)ch[j] * K;                                                    \
                    for (size_t r = 0; r < K; r++) dst[r] += src[r];                                           \
                }                                                                                              \
                free(ch);                                                                                      \
            }                                                                                                  \
            for (size_t r = 0; r < K; r++)                                                                     \
                for (uint32_t t = 0; t < M; t++) {                                                             \
                    const uint32_t m = order[t], node = g->mut_node[m];                                        \
                    if (node != ORACLE_INVALID_NODE) out[r * M + m] += val[(size_t)node * K + r];              \
                }                                                                                              \
        }                                                                                                      \
        free(order);                                                                                           \
        free(val);                                                                                             \
        return ORACLE_OK;                                                                                      \
    }

DEFINE_FAITHFUL(f64, double)
DEFINE_FAITHFUL(f32, float)
DEFINE_FAITHFUL(i64, int64_t)
DEFINE_FAITHFUL(i32, int32_t)

int oracle_matmul_faithful(const oracle_encoded_graph* eg, int dtype, const void* input, size_t input_rows,
                           int direction, void* output) {
    switch (dtype) {
    case ORACLE_F64: return faithful_f64(eg, (const double*)input, input_rows, direction, (double*)output);
    case ORACLE_F32: return faithful_f32(eg, (const float*)input, input_rows, direction, (float*)output);
    case ORACLE_I64: return faithful_i64(eg, (const int64_t*)input, input_rows, direction, (int64_t*)output);
    case ORACLE_I32: return faithful_i32(eg, (const int32_t*)input, input_rows, direction, (int32_t*)output);
    default: return fail(ORACLE_ERR_MISUSE, "faithful mode supports f64/f32/i64/i32 only");
    }
}
