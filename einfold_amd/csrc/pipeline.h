// Host launchers of the fused pipeline's general kernels (pipeline.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.h"

namespace ef {

// One pipeline over placed matrices (device memory): q [nx][ne] and
// kt [ne][nj] for a product producer, or up to 8 q-side [nx][ne] and
// k-side [ne][nj] columns for an expression; v [nj][ny]; each row's allowed
// keys [lo, hi) (global j); the weight's scale or expression.
struct PipeArgs {
    int64_t nx, ne, nj, ny;
    const double *q, *kt, *v;
    const double *q_cols[8];
    const double *k_cols[8];
    int n_q_cols, n_k_cols;
    const ef_expr_op *score;        // device; n_score == 0: the product
    int n_score;
    // A score of the form (q - k)^2 followed by up to four operations with
    // literals (sq_q >= 0): the q-side and k-side column slots, whether the
    // difference is k - q, and the operations (EF_EXPR_NEG, or _MUL, _DIV,
    // _ADD, _SUB with post_value, the literal on the left when
    // post_literal_first).  sq_q = -1: interpret `score`.
    int sq_q, sq_k, sq_swap, n_post;
    uint32_t post_code[4];
    int post_literal_first[4];
    double post_value[4];
    // sq_mono: the score (after the weight's scale) is a monotone function
    // of the square (every literal finite, no product or quotient by 0), so
    // a tile's extreme scores are those of its extreme squares.
    // sq_neg_div: the operations are exactly a negation then a division by
    // post_value[1], which is one division by -post_value[1] (the same bits);
    // sq_div_pow2: that divisor is a power of two whose reciprocal is a
    // normal double, sq_neg_recip = 1 / -post_value[1] exactly, and the
    // division is that product (the same real quotient, rounded once: the
    // same bits).
    int sq_mono, sq_neg_div, sq_div_pow2;
    double sq_neg_recip;
    const ef_expr_op *weight;       // device; n_weight > 0: Weight::Of
    int n_weight;
    uint32_t scale_kind;
    double scale;
    int shifted;
    uint32_t normalizer;
    const int64_t *lo, *hi;
    const uint8_t *finite, *has_inf;
    const double *largest, *smallest;
    // Per 64-key tile of the J positions: the largest and smallest of its
    // keys' largest and smallest magnitudes, and whether every value of
    // every key in it is finite.
    const double *tile_largest, *tile_smallest;
    const uint8_t *tile_finite;
    double *out;
    uint8_t *special;
    int *overflowed;
};

ef_status value_stats(const double *v, int64_t nj, int64_t ny, uint8_t *finite, uint8_t *has_inf, double *largest,
                      double *smallest, hipStream_t s);
// The per-tile summaries of value_stats' per-key arrays ((nj + 63) / 64
// tiles).
ef_status tile_stats(const uint8_t *finite, const double *largest, const double *smallest, int64_t nj,
                     uint8_t *tile_finite, double *tile_largest, double *tile_smallest, hipStream_t s);
// The largest row norm of a matrix (row r, column c at r*row_stride +
// c*col_stride) into *out (a non-negative double, initialized by the
// caller), and whether any entry isn't finite.
ef_status max_norm(const double *m, int64_t rows, int64_t cols, int64_t row_stride, int64_t col_stride,
                   double *out, int *not_finite, hipStream_t s);
// out[0] = max largest[j], out[1] = min smallest[j] (initialized 0 and
// +inf by the caller); not_finite if some key's values aren't.
ef_status reduce_value_stats(const uint8_t *finite, const double *largest, const double *smallest, int64_t nj,
                             double *out, int *not_finite, hipStream_t s);
// The general tile loop, one thread per row, or with parts > 1 each row's
// tiles split into `parts` runs (one thread each) merged in run order:
// scratch of pipeline_parts_bytes(p, parts) bytes.
size_t pipeline_parts_bytes(const PipeArgs &p, int parts);
ef_status pipeline_rows(const PipeArgs &p, int parts, void *scratch, hipStream_t s);
// in [rows][cols] as out [cols][rows].
ef_status transpose_f64(const double *in, int64_t rows, int64_t cols, double *out, hipStream_t s);
ef_status narrow_f32(const double *in, int64_t n, float *out, hipStream_t s);

}  // namespace ef
