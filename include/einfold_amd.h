/*
 * einfold_amd — the C ABI of an MI355X-native kernel layer for einfold's
 * EinFoldExec operator (xqlsystems/einfold, crates/einfold-kernels).
 *
 * Three levels, each a drop-in for one reference interface:
 *
 *  1. The kernel API that EinFoldExec calls
 *     (crates/einfold-datafusion/src/ein_fold_exec.rs:674-727):
 *     ef_fold_dense, ef_fold and ef_pipeline take the fold or pipeline
 *     description (a C mirror of crates/einfold-ir), the inputs as Arrow
 *     record batches through the Arrow C Data Interface, and options, and
 *     return Arrow record batches.  EF_DECLINED is the reference's
 *     `Ok(None)` (the caller runs the host plan); EF_OUT_OF_MEMORY is its
 *     `ArrowError::MemoryError` from a refused Budget, and is also returned
 *     when device memory runs short (the caller frees its reservation and
 *     runs the host plan); every other non-OK status fails the query.
 *
 *  2. The DenseBackend plugin (crates/einfold-kernels/src/backend.rs:124-143):
 *     ef_dense_may_run and ef_dense_run_product / ef_dense_run_attention
 *     take the packed f64 operands einfold hands a backend.
 *
 *  3. Device-resident entry points (pointers into HBM, a HIP stream) for
 *     callers that keep their columns on the GPU: the dense contraction,
 *     radix ranking, the sparse product and fused attention forward and
 *     backward in FP32, BF16x3 (three BF16 MFMA products recovering FP32
 *     accuracy) or BF16.
 *
 * Conventions: no C++ types cross the ABI; errors are statuses, never
 * exceptions; ef_last_error() gives the calling thread's last message.
 * Results are deterministic: a fixed reduction order and no float atomics,
 * so the bits do not depend on batching, arrival order or launch shape.
 * Calls on distinct streams may run concurrently.
 */
#ifndef EINFOLD_AMD_H
#define EINFOLD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Arrow C Data Interface (the published, ABI-stable structs).          */
/* ------------------------------------------------------------------ */
#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE
#define ARROW_FLAG_DICTIONARY_ORDERED 1
#define ARROW_FLAG_NULLABLE 2
#define ARROW_FLAG_MAP_KEYS_SORTED 4
struct ArrowSchema {
    const char *format;
    const char *name;
    const char *metadata;
    int64_t flags;
    int64_t n_children;
    struct ArrowSchema **children;
    struct ArrowSchema *dictionary;
    void (*release)(struct ArrowSchema *);
    void *private_data;
};
struct ArrowArray {
    int64_t length;
    int64_t null_count;
    int64_t offset;
    int64_t n_buffers;
    int64_t n_children;
    const void **buffers;
    struct ArrowArray **children;
    struct ArrowArray *dictionary;
    void (*release)(struct ArrowArray *);
    void *private_data;
};
#endif

/* ------------------------------------------------------------------ */
/* Status                                                               */
/* ------------------------------------------------------------------ */
typedef enum ef_status {
    EF_OK = 0,
    EF_DECLINED = 1,        /* Ok(None): run the host plan */
    EF_OUT_OF_MEMORY = 2,   /* budget refused or device memory short */
    EF_INVALID = 3,         /* the inputs don't fit the description */
    EF_DEVICE_ERROR = 4     /* a HIP call failed */
} ef_status;

/* The calling thread's last error message ("" if none). */
const char *ef_last_error(void);
/* ABI version: major * 10000 + minor * 100 + patch. */
int ef_version(void);
/* Selects the device for the calling thread's later calls. */
ef_status ef_init(int device);

/* ------------------------------------------------------------------ */
/* IR mirror: crates/einfold-ir/src/fold.rs:58-388, pipeline.rs:60-201, */
/* expression.rs:59-135.                                                */
/* ------------------------------------------------------------------ */
typedef enum { EF_EQUAL = 0, EF_NOT_DISTINCT = 1 } ef_key_equality;          /* fold.rs:66-73 */
typedef enum { EF_LESS = 0, EF_LESS_OR_EQUAL = 1, EF_GREATER = 2,
               EF_GREATER_OR_EQUAL = 3 } ef_comparison;                      /* fold.rs:140-152 */

typedef struct ef_key { uint32_t dimension; uint32_t column; } ef_key;      /* fold.rs:84-91 */

typedef struct ef_operand {                                                  /* fold.rs:101-111 */
    const ef_key *keys;
    uint32_t n_keys;
    int32_t factor;         /* column of the factor, -1 for a mask */
} ef_operand;

typedef struct ef_output {                                                   /* fold.rs:224-232 */
    uint32_t dimension;
    int32_t domain_column;  /* -1 without a domain */
    uint32_t domain_equality;
} ef_output;

typedef struct ef_condition {                                                /* fold.rs:190-199 */
    uint32_t left;
    uint32_t comparison;    /* ef_comparison */
    uint32_t right;
    uint32_t pad;
    int64_t offset;         /* left <cmp> right + offset, exact */
} ef_condition;

typedef struct ef_domain {                                                   /* fold.rs:333-339 */
    const uint32_t *carry;
    uint32_t n_carry;
    uint32_t has_missing;
    double missing;
} ef_domain;

/* A row expression in postfix (expression.rs:121-135). */
typedef enum { EF_EXPR_COLUMN = 0, EF_EXPR_LITERAL = 1, EF_EXPR_NEG = 2, EF_EXPR_ABS = 3,
               EF_EXPR_EXP = 4, EF_EXPR_LN = 5, EF_EXPR_ADD = 6, EF_EXPR_SUB = 7,
               EF_EXPR_MUL = 8, EF_EXPR_DIV = 9, EF_EXPR_POW = 10 } ef_expr_code;
typedef struct ef_expr_op {
    uint32_t code;
    uint32_t operand;       /* EF_EXPR_COLUMN: the operand's index */
    uint32_t column;        /* EF_EXPR_COLUMN: its column */
    uint32_t pad;
    double value;           /* EF_EXPR_LITERAL */
} ef_expr_op;

typedef struct ef_fold_desc {                                                     /* fold.rs:380-388 */
    const uint32_t *equality;       /* per dimension, ef_key_equality */
    uint32_t n_dimensions;
    uint32_t n_operands;
    const ef_operand *operands;
    const ef_output *outputs;
    uint32_t n_outputs;
    uint32_t n_conditions;
    const ef_condition *conditions;
    const ef_domain *domain;        /* NULL: none */
    const ef_expr_op *row;          /* NULL: RowValue::Product */
    uint32_t n_row;
    uint32_t pad;
} ef_fold_desc;

typedef enum { EF_SCALE_ONE = 0, EF_SCALE_TIMES = 1, EF_SCALE_OVER = 2 } ef_scale_kind;   /* pipeline.rs:63-70 */
typedef enum { EF_SHIFT_NONE = 0, EF_SHIFT_KEPT_MAXIMUM = 1 } ef_shift;                   /* pipeline.rs:123-136 */
typedef enum { EF_NORM_KEPT_GROUPS = 0, EF_NORM_JOINED_ROWS = 1, EF_NORM_NONE = 2 } ef_normalizer; /* pipeline.rs:141-152 */

typedef struct ef_pipeline_desc {                                                 /* pipeline.rs:193-201 */
    ef_fold_desc producer;
    uint32_t weight_kind;           /* 0: Weight::Exp, 1: Weight::Of */
    uint32_t scale_kind;
    double scale;
    uint32_t shift;
    uint32_t n_weight;
    const ef_expr_op *weight;       /* Weight::Of's expression of the score */
    const uint32_t *equality;       /* the pipeline's own dimensions */
    uint32_t n_dimensions;
    uint32_t n_outputs;
    const uint32_t *outputs;
    ef_operand values;
    uint32_t normalizer;
    uint32_t pad;
} ef_pipeline_desc;

/* One input relation (crates/einfold-kernels/src/relation.rs:13-18): its
 * schema and its record batches (struct arrays), in arrival order. */
typedef struct ef_relation {
    const struct ArrowSchema *schema;
    struct ArrowArray *const *batches;
    size_t n_batches;
} ef_relation;

/* Options (crates/einfold-kernels/src/options.rs:70-121). */
typedef enum { EF_FLOAT_KEYS_BITWISE = 0, EF_FLOAT_KEYS_SIGNED_ZEROS_EQUAL = 1 } ef_float_keys;
typedef enum { EF_DOMAIN_ROWS_ANY = 0, EF_DOMAIN_ROWS_UNIQUE_KEYS = 1 } ef_domain_rows;
typedef enum { EF_PRECISION_FP64 = 0, EF_PRECISION_FP32 = 1, EF_PRECISION_BF16X3 = 2,
               EF_PRECISION_BF16 = 3 } ef_precision;
typedef struct ef_options {
    uint32_t float_keys;            /* ef_float_keys */
    uint32_t threads;               /* accepted; bits never depend on it */
    uint32_t domain_rows;           /* ef_domain_rows */
    /* Arithmetic for Float32 factors (with accept_float32).  EF_PRECISION_FP64:
     * widened to Float64, the reference's exact summation order.  _FP32,
     * _BF16X3, _BF16: the dense fold's contraction on MFMA with FP32
     * accumulation (where every factor is finite), the fused pipeline in
     * that precision. */
    uint32_t precision;
    /* Budget (options.rs:44-66): grow returns nonzero if `bytes` more may
     * be allocated; shrink gives back bytes grow counted.  NULL: no limit. */
    void *budget_ctx;
    int (*budget_grow)(void *ctx, size_t bytes);
    void (*budget_shrink)(void *ctx, size_t bytes);
    void *stream;                   /* hipStream_t; NULL: the default stream */
    uint32_t accept_float32;        /* take Float32 factors (the reference declines them) */
    /* 1: a dense fold of finite Float64 factors may run its contraction on
     * FP64 MFMA, in another summation order, within the reference's bound
     * 2 (n - 1) u sum|terms| instead of its exact bits.  0: exact order. */
    uint32_t fp64_mfma;
} ef_options;

typedef enum { EF_ALGORITHM_DENSE = 0, EF_ALGORITHM_DISCRIMINATE = 1, EF_ALGORITHM_FUSED = 2 } ef_algorithm;

/* A kernel's result: owned Arrow batches of `schema`.  A consumer may move
 * any of them out (importing sets its release to NULL); ef_result_release
 * releases what is left. */
typedef struct ef_result {
    struct ArrowSchema schema;
    struct ArrowArray *batches;
    size_t n_batches;
    uint32_t algorithm;             /* ef_algorithm */
    uint32_t pad;
    uint64_t tiles_visited;         /* fused.rs:113-123 */
    uint64_t tiles;
} ef_result;

void ef_result_release(ef_result *result);

/* ------------------------------------------------------------------ */
/* 1. Kernel API (crates/einfold-kernels/src/algorithm.rs:52-160)        */
/* ------------------------------------------------------------------ */

/* fold_dense (algorithm.rs:52-62, dense.rs:176-578): the dense positional
 * algorithm alone; EF_DECLINED where its requirements don't hold. */
ef_status ef_fold_dense(const ef_fold_desc *fold, const ef_relation *a, const ef_relation *b,
                        const ef_relation *domain, const ef_options *options, ef_result *out);

/* Many dense folds in one call, each as ef_fold_dense would compute it
 * (algorithm.rs:52-62; EinFoldExec calls fold_dense once per fold,
 * ein_fold_exec.rs:674-727): every fold's keys are cast and its operands
 * placed, then every contraction that runs on MFMA (Float32 factors with an
 * MFMA precision) joins one persistent grouped launch, then each fold
 * emits.  a, b: `count` relations each; domains: `count` pointers (NULL
 * for none) or NULL.  status[i] and out[i] are fold i's status and result
 * (EF_DECLINED, EF_OUT_OF_MEMORY and EF_INVALID per fold).  Returns EF_OK
 * unless the device failed. */
ef_status ef_fold_dense_many(size_t count, const ef_fold_desc *const *folds, const ef_relation *a,
                             const ef_relation *b, const ef_relation *const *domains, const ef_options *options,
                             ef_result *out, ef_status *status);

/* fold (algorithm.rs:97-133): dense, else discrimination (radix ranking,
 * counting sorts, Gustavson's product: discriminate.rs:369-650). */
ef_status ef_fold(const ef_fold_desc *fold, const ef_relation *a, const ef_relation *b,
                  const ef_relation *domain, const ef_options *options, ef_result *out);

/* fold_chain (algorithm.rs:76-83, chain.rs:55-288): a SUM of the product
 * of three operands' factors, with no domain and no conditions, as two
 * dense folds, the pair whose two products cost the fewest cells first;
 * the intermediate stays an owned result between them.  EF_DECLINED
 * where that doesn't apply: a non-finite factor, sums that could
 * overflow (the product over the operands of their largest magnitude
 * times their rows above 2^1000), or keys the dense algorithm declines. */
ef_status ef_fold_chain(const ef_fold_desc *fold, const ef_relation *a, const ef_relation *b,
                        const ef_relation *c, const ef_options *options, ef_result *out);

/* pipeline (algorithm.rs:146-160, fused.rs:151-489): the fused tile
 * algorithm with an online softmax and mask tile skipping. */
ef_status ef_pipeline(const ef_pipeline_desc *pipeline, const ef_relation *q, const ef_relation *k,
                      const ef_relation *v, const ef_options *options, ef_result *out);

/* ------------------------------------------------------------------ */
/* 2. DenseBackend (crates/einfold-kernels/src/backend.rs:17-143)        */
/* ------------------------------------------------------------------ */
typedef struct ef_product_shape { size_t batches, rows, inner, columns; } ef_product_shape;        /* backend.rs:17-28 */
typedef struct ef_attention_shape { size_t batches, queries, features, keys, values; } ef_attention_shape; /* backend.rs:30-45 */
typedef enum { EF_WORK_PRODUCT = 0, EF_WORK_ATTENTION = 1 } ef_work_kind;

/* may_run (backend.rs:124-133): whether the device takes work of this
 * kind and shape (it fits device memory and the kernels' index range). */
int ef_dense_may_run(uint32_t kind, const ef_product_shape *product, const ef_attention_shape *attention);
/* run, DenseWork::Product (backend.rs:62-73): out[w][x][y] = a[w][x][z] b[w][z][y]. */
ef_status ef_dense_run_product(const double *a, const double *b, ef_product_shape shape,
                               double *out, const ef_options *options);
/* run, DenseWork::Attention (backend.rs:74-95): shifted softmax of scaled
 * q kt over each query's allowed keys [allowed[2x], allowed[2x+1]). */
ef_status ef_dense_run_attention(const double *q, const double *kt, const double *v,
                                 ef_attention_shape shape, uint32_t scale_kind, double scale,
                                 const size_t *allowed, double *out, const ef_options *options);

/* ------------------------------------------------------------------ */
/* 3. Device-resident entry points (all pointers in device memory)      */
/* ------------------------------------------------------------------ */

/* dense.rs:706-770 in the reference's exact order (bit-identical sums):
 * out[x][y] = sum over z of a[x][z] * b[x / per_batch][z][y] for cells
 * whose held[x][z] == 2 (valued), each product and sum rounded once. */
ef_status ef_dense_accumulate_f64(const double *a, const uint32_t *held, const double *b,
                                  size_t nx, size_t nz, size_t ny, size_t per_batch,
                                  double *out, void *stream);

/* Batched FP32 products out[w] = a[w] (m x k) * b[w] (k x n), row-major,
 * FP32 accumulation; precision EF_PRECISION_FP32, _BF16X3 or _BF16. */
ef_status ef_dense_product_f32(const float *a, const float *b, float *out, size_t batches,
                               size_t m, size_t n, size_t k, uint32_t precision, void *stream);

/* Ranking (discriminate.rs:103-156, 280-292): rank[i] is the number of
 * distinct words below words[i]; first[r] the lowest row holding the r-th
 * distinct word (first holds n entries).  *n_ranks (host) receives the
 * number of ranks.  `scratch` is device memory of ef_rank_scratch_bytes(n)
 * bytes.  From 65536 rows the words are deduplicated by hashing first;
 * the radix sort ranks the rest and what hashing cannot finish. */
size_t ef_rank_scratch_bytes(size_t n);
ef_status ef_rank_u64(const uint64_t *words, size_t n, uint32_t *rank, uint32_t *first,
                      size_t *n_ranks, void *scratch, void *stream);

/* Fused attention over (batch, position, feature) tensors, row-major:
 * q [batches][queries][features], k [batches][keys][features],
 * v [batches][keys][values]; each query x sees keys [lo(x), hi(x)) where
 * lo/hi come from `allowed` (device, 2 per query, the same in every batch)
 * or, if NULL, from causal/window: window <= 0 is causal
 * (key <= query + causal_offset), else also key > query - window.
 * Scores are scaled by `scale` (multiplied).  Writes out
 * [batches][queries][values] and per-row statistics row_max and row_sum
 * [batches][queries] (the softmax's shift and normalizer, for the
 * backward pass): float32 for EF_PRECISION_FP32, _BF16X3 and _BF16,
 * float64 for EF_PRECISION_FP64.  A query with no allowed key gets NaN.
 * precision: EF_PRECISION_FP32, _BF16X3, _BF16 (float32 tensors) or
 * EF_PRECISION_FP64 (double tensors). */
typedef struct ef_attention_desc {
    size_t batches, queries, keys, features, values;
    double scale;
    const int64_t *allowed;         /* device, 2 per query, or NULL */
    int64_t causal;                 /* 1: key <= query + causal_offset */
    int64_t causal_offset;
    int64_t window;                 /* > 0: also key > query + causal_offset - window */
    uint32_t precision;
    uint32_t pad;
} ef_attention_desc;

/* Device scratch for the attention entry points below (the operands
 * prepared in MFMA fragment order, work counters and per-row sums): one
 * buffer of this many bytes serves the forward and backward passes of
 * `desc`.  Calls that share a buffer must run in order (one stream). */
size_t ef_attention_scratch_bytes(const ef_attention_desc *desc);

ef_status ef_attention_fwd(const ef_attention_desc *desc, const void *q, const void *k,
                           const void *v, void *out, void *row_max, void *row_sum,
                           uint64_t *tiles_visited, void *scratch, void *stream);

/* No counterpart in the reference: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  ef_attention_fwd with its scratch in a
 * fixed device array of EF_NOALLOC_SCRATCH_BYTES per GPU (held until the
 * process ends) instead of a caller's buffer; the caller still passes the
 * tensors and outputs.  It allocates nothing.  A description whose forward
 * pass needs more scratch than the array holds is rejected (EF_INVALID,
 * the needed bytes in ef_last_error()); at 8 heads x 4096 queries and keys
 * x 64 features the FP32 and BF16 precisions need about 33 MB.  Every call
 * on one GPU shares its array: calls that overlap on one GPU race. */
#define EF_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
ef_status ef_attention_fwd_noalloc(const ef_attention_desc *desc, const void *q, const void *k,
                                   const void *v, void *out, void *row_max, void *row_sum,
                                   uint64_t *tiles_visited, void *stream);

/* The backward pass of ef_attention_fwd, recomputing score tiles from the
 * saved row_max/row_sum (never storing the T x T scores): given dout
 * (the upstream gradient), writes dq, dk, dv.  32 or 64 features. */
ef_status ef_attention_bwd(const ef_attention_desc *desc, const void *q, const void *k,
                           const void *v, const void *out, const void *dout,
                           const void *row_max, const void *row_sum,
                           void *dq, void *dk, void *dv, void *scratch, void *stream);

/* ef_attention_bwd in its three stages, for callers that want only some
 * gradients (a frozen query table needs no dq): prepare fills the scratch
 * (the operands in fragment order and each query's sum of dout * out);
 * bwd_kv then gives dk and dv, bwd_q gives dq, each from that scratch. */
ef_status ef_attention_bwd_prepare(const ef_attention_desc *desc, const void *q, const void *k,
                                   const void *v, const void *out, const void *dout,
                                   void *scratch, void *stream);
ef_status ef_attention_bwd_kv(const ef_attention_desc *desc, const void *k, const void *v,
                              const void *row_max, const void *row_sum, void *scratch,
                              void *dk, void *dv, void *stream);
ef_status ef_attention_bwd_q(const ef_attention_desc *desc, const void *q, const void *dout,
                             const void *row_max, const void *row_sum, void *scratch,
                             void *dq, void *stream);

/* Many uneven FP32 products in one persistent launch (batched dense folds
 * with Float32 factors, dense.rs:706-770): c = a b per problem, row-major
 * a [m][k], b [k][n], c [m][n], device pointers.  A plan prepares the
 * problems once (work list, operand buffers, the tile shape its autotuner
 * measured fastest for this set of shapes unless `tile` is 64, 128 or, for
 * the BF16 precisions, 256: 256 x 256 tiles for BF16, 256 x 128 for
 * BF16x3, for large products) and
 * captures prep + GEMM as a HIP graph; each run replays it on `stream`.
 * The plan reads the problems' pointers at every run and must be rebuilt
 * when they change.  precision: EF_PRECISION_FP32, _BF16X3 or _BF16. */
typedef struct ef_gemm_problem {
    const float *a;
    const float *b;
    float *c;
    size_t m, n, k;
} ef_gemm_problem;
typedef struct ef_grouped_plan ef_grouped_plan;
ef_status ef_grouped_plan_create(const ef_gemm_problem *problems, size_t count, uint32_t precision, int tile,
                                 void *stream, ef_grouped_plan **plan);
ef_status ef_grouped_plan_run(ef_grouped_plan *plan, void *stream);
/* The plan's tile edge in rows (64, 128 or 256). */
int ef_grouped_plan_tile(const ef_grouped_plan *plan);
void ef_grouped_plan_destroy(ef_grouped_plan *plan);

#ifdef __cplusplus
}
#endif
#endif /* EINFOLD_AMD_H */
