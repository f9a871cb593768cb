// Host launchers of the discrimination algorithm's device kernels
// (discriminate.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "arrow_io.h"
#include "common.h"

namespace ef {

// Gustavson's product over sorted rows: a's rows in (x, z, factor) order
// with their z ranks and factors; b's in (z, y, factor) order with their y
// ranks and factors; a_start per x rank, b_start per z rank.  Each wave
// slot owns acc[ny] (zero) and state[ny] (zero) and returns them zeroed.
// Groups of x go to [bound_off[x], bound_off[x] + g_count[x]) in y order.
struct GustavsonArgs {
    const uint32_t *a_start, *a_z;
    const uint8_t *a_valid;
    const double *a_value;
    const uint32_t *b_start, *b_y;
    const uint8_t *b_valid;
    const double *b_value;
    int64_t nx, ny;
    double *acc;
    uint8_t *state;
    unsigned long long *next;
    const uint64_t *bound_off;
    uint32_t *g_y;
    double *g_sum;
    uint8_t *g_valued;
    uint32_t *g_count;
};

// The output's key columns: each taken from a's (from_a) or b's key column
// at the first row of the group's rank.
struct GroupKeys {
    const void *src[8];
    const uint8_t *src_valid[8];
    void *dst[8];
    uint8_t *dst_valid[8];
    int width[8];
    int from_a[8];
    int n;
};

size_t scan64_scratch_bytes(size_t n);
ef_status scan_u64(const uint64_t *in, uint64_t *out, size_t n, void *tmp, hipStream_t s);
ef_status order_words(Type t, const void *raw, const uint8_t *valid, int64_t n, uint64_t *out, hipStream_t s);
ef_status zero_signs(Type t, void *raw, int64_t n, hipStream_t s);
ef_status null_codes(const uint32_t *rank, const uint8_t *valid, int64_t n, uint64_t *out, hipStream_t s);
ef_status combine_ranks(const uint32_t *hi, const uint32_t *lo, int64_t n, uint64_t *out, hipStream_t s);
ef_status and_valid(uint8_t *join, const uint8_t *valid, int64_t n, hipStream_t s);
ef_status select_rows(const uint8_t *keep, int64_t n, uint32_t *flags, uint32_t *excl, uint32_t *rows,
                      void *scan_tmp, uint32_t *count, hipStream_t s);
ef_status pair_words(const uint32_t *rows, int64_t n, const uint32_t *hi, const uint32_t *lo, uint64_t *out,
                     hipStream_t s);
ef_status tie_words(const uint32_t *rows, int64_t n, const uint8_t *valid, const double *value, uint64_t *bits,
                    uint64_t *valid_w, hipStream_t s);
ef_status any_repeated(const uint64_t *w, int64_t n, int *flag, bool *out, hipStream_t s);
ef_status segment_starts(const uint64_t *w, int64_t n, int64_t nk, uint32_t *start, hipStream_t s);
ef_status gather_sorted(const uint32_t *rows, const uint64_t *w, int64_t n, const uint8_t *valid,
                        const double *value, uint32_t *low, uint8_t *valid_s, double *value_s, hipStream_t s);
ef_status group_bounds(const uint32_t *a_start, const uint32_t *a_z, const uint32_t *b_start, int64_t nx, int64_t ny,
                       uint64_t *bound, unsigned long long *pairs, hipStream_t s);
// Gustavson's product for operands with heavy join ranks: g.acc and
// g.state hold one zeroed cell per (x, y) (nx * ny), heavy_host/heavy_dev
// the heavy z ranks in ascending order (host and device copies),
// heavy_rows their b bucket sizes; scratch: split 2 * n_heavy * nx words,
// runs one word per heavy bucket row (heavy z k's from run_off[k], host
// and device copies of the offsets), n_runs n_heavy words.  Leaves each
// x's groups staged as gustavson does.
ef_status gustavson_heavy(const GustavsonArgs &g, const uint32_t *heavy_host, const uint32_t *heavy_dev,
                          const uint32_t *heavy_rows, int n_heavy, uint32_t *split, uint32_t *runs,
                          const uint64_t *run_off_host, const uint64_t *run_off_dev, uint32_t *n_runs, hipStream_t s);
ef_status gustavson(const GustavsonArgs &g, int blocks, hipStream_t s);
ef_status widen_counts(const uint32_t *in, int64_t n, uint64_t *out, hipStream_t s);
// Device bytes emit_groups needs in `scratch` for ny y ranks.
size_t emit_scratch_bytes(int64_t ny);
// The groups as output rows (keys, sums, validity); ny: the y ranks that
// exist (y_first[0 .. ny - 1] valid).
ef_status emit_groups(int64_t nx, int64_t ny, const uint64_t *bound_off, const uint64_t *out_off,
                      const uint32_t *g_count, const uint32_t *g_y, const double *g_sum, const uint8_t *g_valued,
                      const uint32_t *x_first, const uint32_t *y_first, const GroupKeys &keys, void *scratch,
                      double *values, uint8_t *valid, hipStream_t s);
ef_status domain_lookup(int64_t n, const uint8_t *matches, const uint32_t *x_rank, const uint32_t *y_rank, int64_t nx,
                        const uint64_t *bound_off, const uint32_t *g_count, const uint32_t *g_y, const double *g_sum,
                        const uint8_t *g_valued, int has_missing, double missing, double *values, uint8_t *valid,
                        hipStream_t s);

}  // namespace ef
