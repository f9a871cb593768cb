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
    const ef_expr_op *weight;       // device; n_weight > 0: Weight::Of
    int n_weight;
    uint32_t scale_kind;
    double scale;
    int shifted;
    uint32_t normalizer;
    const int64_t *lo, *hi;
    const uint8_t *finite, *has_inf;
    const double *largest, *smallest;
    double *out;
    uint8_t *special;
    int *overflowed;
};

ef_status value_stats(const double *v, int64_t nj, int64_t ny, uint8_t *finite, uint8_t *has_inf, double *largest,
                      double *smallest, hipStream_t s);
// The largest row norm of a matrix (row r, column c at r*row_stride +
// c*col_stride) into *out (a non-negative double, initialized by the
// caller), and whether any entry isn't finite.
ef_status max_norm(const double *m, int64_t rows, int64_t cols, int64_t row_stride, int64_t col_stride,
                   double *out, int *not_finite, hipStream_t s);
// out[0] = max largest[j], out[1] = min smallest[j] (initialized 0 and
// +inf by the caller); not_finite if some key's values aren't.
ef_status reduce_value_stats(const uint8_t *finite, const double *largest, const double *smallest, int64_t nj,
                             double *out, int *not_finite, hipStream_t s);
ef_status pipeline_rows(const PipeArgs &p, hipStream_t s);
// in [rows][cols] as out [cols][rows].
ef_status transpose_f64(const double *in, int64_t rows, int64_t cols, double *out, hipStream_t s);
ef_status narrow_f32(const double *in, int64_t n, float *out, hipStream_t s);

}  // namespace ef
