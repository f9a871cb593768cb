// The fused pipeline's device kernels other than the MFMA attention path
// (crates/einfold-kernels/src/fused.rs:371-384, 694-1086): per-key
// statistics of the values, the bounds that decide whether the MFMA path's
// guards can trip, and the general FP64 tile loop, one thread per query
// row, for every pipeline the MFMA path doesn't take (row expressions as
// producers, Weight::Of, unshifted or unnormalized weights, infinite or NaN
// values).  A row's result depends only on its own allowed keys and the
// 64-key tile grid, so rows are independent; each follows fused.rs's steps
// for its own row: scores summed over E in order, the tile's shift, the
// rescale of the running sums, exp, the weighted sum of v, normalization,
// and the overflow and underflow guards (Bounds::hold).
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <cstdint>

#include "common.h"
#include "pipeline.h"

namespace ef {

namespace {

constexpr int JT = 64;

// Per key j: every value finite, largest finite magnitude, smallest
// nonzero finite magnitude, any infinite value (fused.rs:371-384).
__global__ __launch_bounds__(256) void value_stats_kernel(const double *__restrict__ v, int64_t nj, int64_t ny,
                                                          uint8_t *finite, uint8_t *has_inf, double *largest,
                                                          double *smallest) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nj) return;
    int f = 1, inf = 0;
    double l = 0.0, s = INFINITY;
    for (int64_t y = 0; y < ny; y++) {
        const double x = v[j * ny + y];
        if (!isfinite(x)) {
            f = 0;
            if (isinf(x)) inf = 1;
            continue;
        }
        const double m = fabs(x);
        l = fmax(l, m);
        if (m != 0.0) s = fmin(s, m);
    }
    finite[j] = (uint8_t)f;
    has_inf[j] = (uint8_t)inf;
    largest[j] = l;
    smallest[j] = s;
}

// One thread per 64-key tile: the summaries the squared-difference loop
// reads instead of each key's statistics when a row covers the whole tile.
__global__ __launch_bounds__(256) void tile_stats_kernel(const uint8_t *finite, const double *largest,
                                                         const double *smallest, int64_t nj, uint8_t *tile_finite,
                                                         double *tile_largest, double *tile_smallest) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t * 64 >= nj) return;
    const int64_t e = (t + 1) * 64 < nj ? (t + 1) * 64 : nj;
    int f = 1;
    double l = 0.0, s = INFINITY;
    for (int64_t j = t * 64; j < e; j++) {
        f &= finite[j];
        l = fmax(l, largest[j]);
        s = fmin(s, smallest[j]);
    }
    tile_finite[t] = (uint8_t)f;
    tile_largest[t] = l;
    tile_smallest[t] = s;
}

__device__ __forceinline__ void atomic_max_pos(double *addr, double v) {
    // Non-negative doubles order as their bits.
    atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

__device__ __forceinline__ void atomic_min_pos(double *addr, double v) {
    atomicMin(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

// Row norms of a [rows][cols] matrix: the largest, and whether every entry
// is finite.  For the transposed kt [cols][rows] layout, stride swaps.
__global__ __launch_bounds__(256) void max_norm_kernel(const double *__restrict__ m, int64_t rows, int64_t cols,
                                                       int64_t row_stride, int64_t col_stride, double *max_norm,
                                                       int *not_finite) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double ss = 0.0;
    int bad = 0;
    if (r < rows) {
        for (int64_t c = 0; c < cols; c++) {
            const double x = m[r * row_stride + c * col_stride];
            if (!isfinite(x)) bad = 1;
            ss += x * x;
        }
    }
    double n = sqrt(ss);
    if (!isfinite(n)) bad = 1;
    for (int o = 32; o >= 1; o >>= 1) {
        n = fmax(n, __shfl_xor(n, o));
        bad |= __shfl_xor(bad, o);
    }
    if ((threadIdx.x & 63) == 0) {
        if (bad) atomicOr(not_finite, 1);
        else atomic_max_pos(max_norm, n);
    }
}

__global__ __launch_bounds__(256) void reduce_stats_kernel(const uint8_t *finite, const double *largest,
                                                           const double *smallest, int64_t nj, double *out,
                                                           int *not_finite) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0, s = INFINITY;
    int bad = 0;
    if (j < nj) {
        l = largest[j];
        s = smallest[j];
        bad = !finite[j];
    }
    for (int o = 32; o >= 1; o >>= 1) {
        l = fmax(l, __shfl_xor(l, o));
        s = fmin(s, __shfl_xor(s, o));
        bad |= __shfl_xor(bad, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomic_max_pos(&out[0], l);
        atomic_min_pos(&out[1], s);
        if (bad) atomicOr(not_finite, 1);
    }
}

// ---- row expressions (crates/einfold-ir/src/expression.rs:59-100) ------

__device__ double eval(const ef_expr_op *e, int n, const double *qv, const double *kv, double score) {
    double st[16];
    int sp = 0;
    for (int i = 0; i < n; i++) {
        const ef_expr_op &op = e[i];
        switch (op.code) {
        case EF_EXPR_COLUMN: st[sp++] = qv ? (op.operand == 0 ? qv[op.column] : kv[op.column]) : score; break;
        case EF_EXPR_LITERAL: st[sp++] = op.value; break;
        case EF_EXPR_NEG: st[sp - 1] = -st[sp - 1]; break;
        case EF_EXPR_ABS: st[sp - 1] = fabs(st[sp - 1]); break;
        case EF_EXPR_EXP: st[sp - 1] = exp(st[sp - 1]); break;
        case EF_EXPR_LN: st[sp - 1] = log(st[sp - 1]); break;
        default: {
            const double b = st[--sp], a = st[sp - 1];
            double r;
            switch (op.code) {
            case EF_EXPR_ADD: r = a + b; break;
            case EF_EXPR_SUB: r = a - b; break;
            case EF_EXPR_MUL: r = a * b; break;
            case EF_EXPR_DIV: r = a / b; break;
            // A square as one correctly rounded product (within pow's own
            // error bound), not the general pow's logarithm and exponential.
            default: r = b == 2.0 ? a * a : pow(a, b); break;
            }
            st[sp - 1] = r;
        }
        }
    }
    return st[0];
}

// fused.rs:1024-1028
__device__ bool overflows(double z, const double *a, bool special_weight, const uint8_t *special_value, int64_t ny) {
    if (special_weight) return false;
    if (!isfinite(z)) return true;
    for (int64_t y = 0; y < ny; y++)
        if (!special_value[y] && !isfinite(a[y])) return true;
    return false;
}

// fused.rs:1058-1086: Bounds::hold.
__device__ bool bounds_hold(double heaviest, double lightest, double largest, double smallest, uint64_t terms_n,
                            bool has_z, double z, const double *results, const uint8_t *special, int64_t n) {
    const double terms = 2.0 * (double)terms_n;
    bool finite = isfinite(heaviest * largest * terms);
    if (finite && has_z) {
        const double quotient = heaviest / fabs(z);
        finite = isfinite(quotient) && isfinite(quotient * largest * terms);
    }
    const double inverse = has_z ? 1.0 / fabs(z) : 1.0;
    const double tiny = __longlong_as_double(1LL);
    const double lost = terms * tiny * (inverse + largest + 1.0);
    const double rounding = DBL_EPSILON / 4.0;
    const double normal = 2.0 * DBL_MIN;
    bool no_underflow = smallest == INFINITY;
    if (!no_underflow) {
        no_underflow = lightest * smallest >= normal;
        if (no_underflow && has_z) {
            const double quotient = lightest / fabs(z);
            no_underflow = quotient >= normal && quotient * smallest >= normal;
        }
    }
    if (!finite) return false;
    if (no_underflow) return true;
    for (int64_t i = 0; i < n; i++)
        if (!special[i] && !(lost <= rounding * fabs(results[i]))) return false;
    return true;
}

__device__ __forceinline__ double scale_apply(const PipeArgs &p, double s) {
    switch (p.scale_kind) {
    case EF_SCALE_TIMES: return s * p.scale;
    case EF_SCALE_OVER: return s / p.scale;
    default: return s;
    }
}

#pragma clang fp contract(off)
// One row's running state over a range of its J tiles (fused.rs:694-935):
// the shift m, the weights' sum z, the guards' extremes and counts.
struct RowState {
    double m, z, high, light, row_largest, row_smallest, low;
    uint64_t terms;
    int special_weight, over;
};

__device__ __forceinline__ void state_init(const PipeArgs &p, RowState &st) {
    const bool shifted = p.n_weight == 0 && p.shifted;
    st.m = shifted ? -INFINITY : 0.0;
    st.z = 0.0;
    st.high = -INFINITY;
    st.light = INFINITY;
    st.row_largest = 0.0;
    st.row_smallest = INFINITY;
    st.low = INFINITY;
    st.terms = 0;
    st.special_weight = 0;
    st.over = 0;
}

// The producer's score of pair (x, j): the product, the interpreted row
// expression, or (FORM 2) a squared difference of one q-side and one
// k-side column followed by up to four operations with literals -- the
// same operations in the same order as the interpreter, so the same bits.
template <int FORM>
__device__ __forceinline__ double score_of(const PipeArgs &p, int64_t x, int64_t j, double *qv, double *kv) {
    const int64_t ne = p.ne, nj = p.nj;
    double s = 0.0;
    if constexpr (FORM == 0) {
        for (int64_t e = 0; e < ne; e++) s += p.q[x * ne + e] * p.kt[e * nj + j];
    } else if constexpr (FORM == 1) {
        for (int64_t e = 0; e < ne; e++) {
            for (int c = 0; c < p.n_q_cols; c++) qv[c] = p.q_cols[c][x * ne + e];
            for (int c = 0; c < p.n_k_cols; c++) kv[c] = p.k_cols[c][e * nj + j];
            const double t = eval(p.score, p.n_score, qv, kv, 0.0);
            s = e == 0 ? t : s + t;
        }
    } else {
        for (int64_t e = 0; e < ne; e++) {
            const double qx = p.q_cols[p.sq_q][x * ne + e], ky = p.k_cols[p.sq_k][e * nj + j];
            const double d = p.sq_swap ? ky - qx : qx - ky;
            double t = d * d;
            for (int i = 0; i < p.n_post; i++) {
                const double c = p.post_value[i];
                switch (p.post_code[i]) {
                case EF_EXPR_NEG: t = -t; break;
                case EF_EXPR_MUL: t = p.post_literal_first[i] ? c * t : t * c; break;
                case EF_EXPR_DIV: t = p.post_literal_first[i] ? c / t : t / c; break;
                case EF_EXPR_ADD: t = p.post_literal_first[i] ? c + t : t + c; break;
                default: t = p.post_literal_first[i] ? c - t : t - c; break;   // EF_EXPR_SUB
                }
            }
            s = e == 0 ? t : s + t;
        }
    }
    return s;
}

// The operations after the square.
__device__ __forceinline__ double sq_post(const PipeArgs &p, double t) {
    if (p.sq_neg_div) return p.sq_div_pow2 ? t * p.sq_neg_recip : t / -p.post_value[1];   // -(t) / c, exactly
    for (int i = 0; i < p.n_post; i++) {
        const double c = p.post_value[i];
        switch (p.post_code[i]) {
        case EF_EXPR_NEG: t = -t; break;
        case EF_EXPR_MUL: t = t * c; break;
        case EF_EXPR_DIV: t = t / c; break;
        case EF_EXPR_ADD: t = t + c; break;
        default: t = t - c; break;   // EF_EXPR_SUB
        }
    }
    return t;
}

__device__ __forceinline__ double sq_score(const PipeArgs &p, double qx, double ky) {
    const double d = p.sq_swap ? ky - qx : qx - ky;
    return sq_post(p, d * d);
}

// fused.rs:1024-1028 for one value column.
__device__ __forceinline__ bool overflows1(double z, double a, bool special_weight, bool special_value) {
    return !special_weight && (!isfinite(z) || (!special_value && !isfinite(a)));
}

// row_range for a squared-difference score over one summed position
// (ne = 1), Weight::Exp and one value column: the running sum in a
// register, each tile's scores computed twice (once for the shift, once for
// the weights) instead of kept in memory, and the tile's key summaries
// read for a tile the row covers whole.  The same operations in the same
// order as row_range, so the same bits.
__device__ void row_range_sq(const PipeArgs &p, int64_t x, int64_t jb, int64_t je, RowState &st, double &acc,
                             uint8_t &special_value) {
    const int64_t nj = p.nj;
    const bool shifted = p.shifted;
    const bool unnormalized = p.normalizer == EF_NORM_NONE;
    const int64_t lo_x = p.lo[x] > jb ? p.lo[x] : jb, hi_x = p.hi[x] < je ? p.hi[x] : je;
    if (lo_x >= hi_x) return;
    const double qx = p.q_cols[p.sq_q][x];
    const double *K = p.k_cols[p.sq_k];
    for (int64_t j0 = lo_x - lo_x % JT; j0 < hi_x; j0 += JT) {
        const int64_t width = JT < nj - j0 ? JT : nj - j0;
        const int64_t b = (lo_x < j0 ? j0 : lo_x), e = (hi_x > j0 + width ? j0 + width : hi_x);
        if (b >= e) continue;
        const int64_t t = j0 / JT;
        const bool whole = b == j0 && e == j0 + width;
        const bool all_finite = p.tile_finite[t] != 0;
        st.terms += (uint64_t)(e - b);
        if (whole) {
            st.row_largest = fmax(st.row_largest, p.tile_largest[t]);
            st.row_smallest = fmin(st.row_smallest, p.tile_smallest[t]);
        } else {
            for (int64_t j = b; j < e; j++) {
                st.row_largest = fmax(st.row_largest, p.largest[j]);
                st.row_smallest = fmin(st.row_smallest, p.smallest[j]);
            }
        }
        double shift = shifted ? st.m : 0.0;
        bool exact = false;
        if (p.sq_mono && all_finite) {
            // The squares' extremes: every rounding step after the square is
            // monotone, so the scores' largest and smallest are the scores of
            // these two (exact, as the max over the keys in order gives the
            // same value), and when both are finite so is every score between.
            double tmin = INFINITY, tmax = -INFINITY;
            bool nan = false;
            for (int64_t j = b; j < e; j++) {
                const double d = p.sq_swap ? K[j] - qx : qx - K[j];
                const double t = d * d;
                nan |= isnan(t);
                tmin = fmin(tmin, t);
                tmax = fmax(tmax, t);
            }
            const double s0 = scale_apply(p, sq_post(p, tmin)), s1 = scale_apply(p, sq_post(p, tmax));
            if (!nan && isfinite(s0) && isfinite(s1)) {
                const double hi_s = fmax(s0, s1), lo_s = fmin(s0, s1);
                st.high = fmax(st.high, hi_s);
                st.light = fmin(st.light, lo_s);
                if (shifted && hi_s > shift) shift = hi_s;
                exact = true;
            }
        }
        if (!exact) {
            for (int64_t j = b; j < e; j++) {
                const double s = scale_apply(p, sq_score(p, qx, K[j]));
                if (isfinite(s)) {
                    st.high = fmax(st.high, s);
                    st.light = fmin(st.light, s);
                }
                if (!all_finite && !p.finite[j] && p.has_inf[j] && isfinite(s)) st.low = fmin(st.low, s);
                if (shifted && (s > shift || isnan(s))) shift = s;
            }
        }
        if (shift == -INFINITY) {
            // Every weight 0 (the tile loop still adds 0 * v, NaN for an
            // infinite value).
            for (int64_t j = b; j < e; j++) acc += 0.0 * p.v[j];
            continue;
        }
        // The running sums rescaled first (their overflow check reads them
        // before this tile's weights), then each weight added to the
        // weights' sum and, times its value, to the weighted sum, in key
        // order: the same sequence of roundings on each sum.
        const double rescale = exp(st.m - shift);
        if (shifted && rescale != 1.0) {
            if (overflows1(st.z, acc, st.special_weight, special_value)) st.over = 1;
            st.z *= rescale;
            acc *= rescale;
        }
        double sum = 0.0;
        bool special_weights = false;
        for (int64_t j = b; j < e; j++) {
            const double w = exp(scale_apply(p, sq_score(p, qx, K[j])) - shift);
            sum += w;
            special_weights |= !isfinite(w);
            acc += w * p.v[j];
        }
        st.special_weight |= special_weights;
        bool special_values = false;
        if (!all_finite)
            for (int64_t j = b; j < e; j++) {
                if (p.finite[j]) continue;
                special_values = true;
                if (!isfinite(p.v[j])) special_value = 1;
            }
        if (unnormalized && (special_weights || special_values)) st.over = 1;
        st.z += sum;
        st.m = shift;
    }
}

// The row's tiles from jb (a tile boundary) up to je, clipped to the row's
// allowed keys, in order: scores, shift, rescale, exp and the weighted sum
// a[y] (fused.rs:771-935).  a and special_value hold the range's own sums.
template <int FORM>
__device__ void row_range(const PipeArgs &p, int64_t x, int64_t jb, int64_t je, RowState &st, double *a,
                          uint8_t *special_value) {
    const int64_t ny = p.ny, nj = p.nj;
    const bool mapped = p.n_weight > 0;
    const bool shifted = !mapped && p.shifted;
    const bool unnormalized = p.normalizer == EF_NORM_NONE;
    const int64_t lo_x = p.lo[x] > jb ? p.lo[x] : jb, hi_x = p.hi[x] < je ? p.hi[x] : je;
    double sc[JT];
    double qv[8], kv[8];
    if (lo_x >= hi_x) return;
    for (int64_t j0 = lo_x - lo_x % JT; j0 < hi_x; j0 += JT) {
        const int64_t width = JT < nj - j0 ? JT : nj - j0;
        const int64_t lo = (lo_x < j0 ? j0 : lo_x) - j0;
        const int64_t hi = (hi_x > j0 + width ? j0 + width : hi_x) - j0;
        if (lo >= hi) continue;
        const int n = (int)(hi - lo);
        for (int i = 0; i < n; i++) sc[i] = score_of<FORM>(p, x, j0 + lo + i, qv, kv);
        st.terms += (uint64_t)n;
        for (int i = 0; i < n; i++) {
            st.row_largest = fmax(st.row_largest, p.largest[j0 + lo + i]);
            st.row_smallest = fmin(st.row_smallest, p.smallest[j0 + lo + i]);
        }
        if (mapped) {
            for (int i = 0; i < n; i++) sc[i] = eval(p.weight, p.n_weight, nullptr, nullptr, sc[i]);
            for (int i = 0; i < n; i++)
                if (!isfinite(sc[i]) || !p.finite[j0 + lo + i]) st.over = 1;
            for (int i = 0; i < n; i++) {
                const double w = fabs(sc[i]);
                st.high = fmax(st.high, w);
                if (w != 0.0) st.light = fmin(st.light, w);
            }
            double sum = -0.0;   // Rust's f64 Sum starts from -0.0
            for (int i = 0; i < n; i++) sum += sc[i];
            st.z += sum;
        } else {
            for (int i = 0; i < n; i++) {
                sc[i] = scale_apply(p, sc[i]);
                if (isfinite(sc[i])) {
                    st.high = fmax(st.high, sc[i]);
                    st.light = fmin(st.light, sc[i]);
                }
            }
            for (int i = 0; i < n; i++) {
                const int64_t j = j0 + lo + i;
                if (!p.finite[j] && p.has_inf[j] && isfinite(sc[i])) st.low = fmin(st.low, sc[i]);
            }
            double shift = 0.0;
            if (shifted) {
                shift = st.m;
                for (int i = 0; i < n; i++)
                    if (sc[i] > shift || isnan(sc[i])) shift = sc[i];
            }
            if (shift == -INFINITY) {
                for (int i = 0; i < n; i++) sc[i] = 0.0;
            } else {
                const double rescale = exp(st.m - shift);
                double sum = 0.0;
                bool special_weights = false;
                for (int i = 0; i < n; i++) {
                    sc[i] = exp(sc[i] - shift);
                    sum += sc[i];
                    if (!isfinite(sc[i])) special_weights = true;
                }
                if (shifted && rescale != 1.0) {
                    if (overflows(st.z, a, st.special_weight, special_value, ny)) st.over = 1;
                    st.z *= rescale;
                    for (int64_t y = 0; y < ny; y++) a[y] *= rescale;
                }
                st.special_weight |= special_weights;
                bool special_values = false;
                for (int i = 0; i < n; i++) {
                    const int64_t j = j0 + lo + i;
                    if (p.finite[j]) continue;
                    special_values = true;
                    for (int64_t y = 0; y < ny; y++)
                        if (!isfinite(p.v[j * ny + y])) special_value[y] = 1;
                }
                if (unnormalized && (special_weights || special_values)) st.over = 1;
                st.z += sum;
                st.m = shift;
            }
        }
        // a += w . v (fused.rs:910-935).
        for (int i = 0; i < n; i++) {
            const double w = sc[i];
            const double *vr = p.v + (j0 + lo + i) * ny;
            for (int64_t y = 0; y < ny; y++) a[y] += w * vr[y];
        }
    }
}

// Normalization and the underflow and overflow guards (fused.rs:941-1013)
// of a row's whole state; a is out's row.
__device__ void row_finish(const PipeArgs &p, RowState &st, double *a, const uint8_t *special_value) {
    const int64_t ny = p.ny;
    const bool mapped = p.n_weight > 0;
    const bool shifted = !mapped && p.shifted;
    const bool unnormalized = p.normalizer == EF_NORM_NONE;
    if (unnormalized) {
        if (shifted && st.m == -INFINITY)
            for (int64_t y = 0; y < ny; y++) a[y] = NAN;
    } else {
        if (overflows(st.z, a, st.special_weight, special_value, ny)) st.over = 1;
        for (int64_t y = 0; y < ny; y++) a[y] /= st.z;
        if (st.z != 0.0 && overflows(st.z, a, st.special_weight, special_value, ny)) st.over = 1;
    }
    const double heaviest = mapped ? st.high : (shifted ? exp(st.high - st.m) : exp(st.high));
    if (!(st.special_weight || st.high == -INFINITY || heaviest == 0.0)) {
        const double lightest = mapped ? st.light : (shifted ? exp(st.light - st.m) : exp(st.light));
        if (!bounds_hold(heaviest, lightest, st.row_largest, st.row_smallest, st.terms, !unnormalized, st.z, a,
                         special_value, ny))
            st.over = 1;
        if (st.low != INFINITY) {
            const double lw = shifted ? exp(st.low - st.m) : exp(st.low);
            const double normal = 2.0 * DBL_MIN;
            if (!(lw >= normal && (unnormalized || lw / fabs(st.z) >= normal))) st.over = 1;
        }
    }
    if (st.over) atomicOr(p.overflowed, 1);
}

// One query row per thread over all its J tiles in order.  out[x][y] is
// the row's running sum a; special[x][y] marks values that were infinite
// or NaN.
template <int FORM>
__global__ __launch_bounds__(64) void pipeline_rows_kernel(PipeArgs p) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= p.nx) return;
    double *a = p.out + x * p.ny;
    uint8_t *special_value = p.special + x * p.ny;
    for (int64_t y = 0; y < p.ny; y++) {
        a[y] = 0.0;
        special_value[y] = 0;
    }
    RowState st;
    state_init(p, st);
    if (FORM == 2 && p.ne == 1 && p.ny == 1 && p.n_weight == 0) {
        double acc = 0.0;
        uint8_t sv = 0;
        row_range_sq(p, x, 0, p.nj, st, acc, sv);
        a[0] = acc;
        special_value[0] = sv;
    } else {
        row_range<FORM>(p, x, 0, p.nj, st, a, special_value);
    }
    row_finish(p, st, a, special_value);
}

// Part `t / nx` of row `t % nx`: a contiguous run of the row's J
// tiles (the row's tiles split into `parts` nearly equal runs), its state
// and sums written to the part's slots.
template <int FORM>
__global__ __launch_bounds__(256) void pipeline_parts_kernel(PipeArgs p, int parts, RowState *states, double *pa,
                                                             uint8_t *psv) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.nx * parts) return;
    // Neighbouring threads take the same part of neighbouring rows: their
    // runs cover the same keys (for rows that see the same keys), so a
    // wave's key loads read one line instead of one per lane.
    const int64_t x = t % p.nx;
    const int part = (int)(t / p.nx);
    double *a = pa + t * p.ny;
    uint8_t *special_value = psv + t * p.ny;
    for (int64_t y = 0; y < p.ny; y++) {
        a[y] = 0.0;
        special_value[y] = 0;
    }
    RowState st;
    state_init(p, st);
    const int64_t lo = p.lo[x], hi = p.hi[x];
    if (lo < hi) {
        const int64_t t0 = lo / JT, tiles = (hi - 1) / JT + 1 - t0;
        const int64_t b = t0 + tiles * part / parts, e = t0 + tiles * (part + 1) / parts;
        if (FORM == 2 && p.ne == 1 && p.ny == 1 && p.n_weight == 0) {
            double acc = 0.0;
            uint8_t sv = 0;
            row_range_sq(p, x, b * JT, e * JT, st, acc, sv);
            a[0] = acc;
            special_value[0] = sv;
        } else {
            row_range<FORM>(p, x, b * JT, e * JT, st, a, special_value);
        }
    }
    states[t] = st;
}

// Each row's parts combined in part order: the shifted sums rescaled to the
// larger shift (each side's sums checked for overflow before it is
// rescaled, as the tile loop checks its running sums), the guards'
// extremes and counts merged; then normalized and guarded as one row.
__global__ __launch_bounds__(256) void pipeline_merge_kernel(PipeArgs p, int parts, const RowState *states,
                                                             const double *pa, const uint8_t *psv) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= p.nx) return;
    const int64_t ny = p.ny;
    const bool shifted = p.n_weight == 0 && p.shifted;
    double *a = p.out + x * ny;
    uint8_t *special_value = p.special + x * ny;
    RowState st;
    state_init(p, st);
    for (int64_t y = 0; y < ny; y++) {
        a[y] = 0.0;
        special_value[y] = 0;
    }
    for (int part = 0; part < parts; part++) {
        const int64_t t = part * p.nx + x;
        const RowState s = states[t];
        const double *ap = pa + t * ny;
        const uint8_t *sp = psv + t * ny;
        st.terms += s.terms;
        st.high = fmax(st.high, s.high);
        st.light = fmin(st.light, s.light);
        st.low = fmin(st.low, s.low);
        st.row_largest = fmax(st.row_largest, s.row_largest);
        st.row_smallest = fmin(st.row_smallest, s.row_smallest);
        st.over |= s.over;
        if (shifted) {
            if (s.m == -INFINITY) continue;   // the part saw no weight
            const double mnew = (st.m == -INFINITY || s.m > st.m || isnan(s.m)) ? s.m : st.m;
            double ra = 1.0, rb = 1.0;
            if (st.m != -INFINITY && mnew != st.m) {
                if (overflows(st.z, a, st.special_weight, special_value, ny)) st.over = 1;
                ra = exp(st.m - mnew);
            }
            if (mnew != s.m) {
                if (overflows(s.z, ap, s.special_weight, sp, ny)) st.over = 1;
                rb = exp(s.m - mnew);
            }
            st.z = st.z * ra + s.z * rb;
            for (int64_t y = 0; y < ny; y++) a[y] = a[y] * ra + ap[y] * rb;
            st.m = mnew;
        } else {
            st.z += s.z;
            for (int64_t y = 0; y < ny; y++) a[y] += ap[y];
        }
        st.special_weight |= s.special_weight;
        for (int64_t y = 0; y < ny; y++) special_value[y] |= sp[y];
    }
    row_finish(p, st, a, special_value);
}
#pragma clang fp contract(on)

// kt [ne][nj] as k [nj][ne] (the attention kernels' key layout), through an
// LDS tile so both sides are coalesced.
__global__ __launch_bounds__(256) void transpose_kernel(const double *__restrict__ in, int64_t rows, int64_t cols,
                                                        double *__restrict__ out) {
    __shared__ double t[32][33];
    const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int i = ty; i < 32; i += 8) {
        const int64_t r = r0 + i, c = c0 + tx;
        if (r < rows && c < cols) t[i][tx] = in[r * cols + c];
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int64_t c = c0 + i, r = r0 + tx;
        if (r < rows && c < cols) out[c * rows + r] = t[tx][i];
    }
}

__global__ __launch_bounds__(256) void narrow_kernel(const double *in, int64_t n, float *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

}  // namespace

ef_status value_stats(const double *v, int64_t nj, int64_t ny, uint8_t *finite, uint8_t *has_inf, double *largest,
                      double *smallest, hipStream_t s) {
    if (nj == 0) return EF_OK;
    hipLaunchKernelGGL(value_stats_kernel, dim3(blocks_for(nj, 256)), dim3(256), 0, s, v, nj, ny, finite, has_inf,
                       largest, smallest);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status tile_stats(const uint8_t *finite, const double *largest, const double *smallest, int64_t nj,
                     uint8_t *tile_finite, double *tile_largest, double *tile_smallest, hipStream_t s) {
    if (nj == 0) return EF_OK;
    const int64_t tiles = (nj + 63) / 64;
    hipLaunchKernelGGL(tile_stats_kernel, dim3(blocks_for(tiles, 256)), dim3(256), 0, s, finite, largest, smallest, nj,
                       tile_finite, tile_largest, tile_smallest);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status max_norm(const double *m, int64_t rows, int64_t cols, int64_t row_stride, int64_t col_stride,
                   double *out, int *not_finite, hipStream_t s) {
    if (rows == 0) return EF_OK;
    hipLaunchKernelGGL(max_norm_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, s, m, rows, cols, row_stride,
                       col_stride, out, not_finite);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status reduce_value_stats(const uint8_t *finite, const double *largest, const double *smallest, int64_t nj,
                             double *out, int *not_finite, hipStream_t s) {
    if (nj == 0) return EF_OK;
    hipLaunchKernelGGL(reduce_stats_kernel, dim3(blocks_for(nj, 256)), dim3(256), 0, s, finite, largest, smallest, nj,
                       out, not_finite);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

size_t pipeline_parts_bytes(const PipeArgs &p, int parts) {
    if (parts <= 1) return 0;
    const size_t n = (size_t)p.nx * (size_t)parts;
    return n * (sizeof(RowState) + 8 * (size_t)p.ny + (size_t)p.ny) + 64;
}

ef_status pipeline_rows(const PipeArgs &p, int parts, void *scratch, hipStream_t s) {
    if (p.nx == 0) return EF_OK;
    const int form = p.n_score == 0 ? 0 : (p.sq_q >= 0 ? 2 : 1);
    if (parts <= 1) {
        switch (form) {
        case 0: hipLaunchKernelGGL(pipeline_rows_kernel<0>, dim3(blocks_for(p.nx, 64)), dim3(64), 0, s, p); break;
        case 1: hipLaunchKernelGGL(pipeline_rows_kernel<1>, dim3(blocks_for(p.nx, 64)), dim3(64), 0, s, p); break;
        default: hipLaunchKernelGGL(pipeline_rows_kernel<2>, dim3(blocks_for(p.nx, 64)), dim3(64), 0, s, p); break;
        }
        EF_HIP(hipGetLastError());
        return EF_OK;
    }
    const size_t n = (size_t)p.nx * (size_t)parts;
    RowState *states = static_cast<RowState *>(scratch);
    double *pa = reinterpret_cast<double *>(states + n);
    uint8_t *psv = reinterpret_cast<uint8_t *>(pa + n * (size_t)p.ny);
    switch (form) {
    case 0: hipLaunchKernelGGL(pipeline_parts_kernel<0>, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, parts, states, pa, psv); break;
    case 1: hipLaunchKernelGGL(pipeline_parts_kernel<1>, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, parts, states, pa, psv); break;
    default: hipLaunchKernelGGL(pipeline_parts_kernel<2>, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, parts, states, pa, psv); break;
    }
    hipLaunchKernelGGL(pipeline_merge_kernel, dim3(blocks_for(p.nx, 256)), dim3(256), 0, s, p, parts,
                       (const RowState *)states, (const double *)pa, (const uint8_t *)psv);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status transpose_f64(const double *in, int64_t rows, int64_t cols, double *out, hipStream_t s) {
    if (rows == 0 || cols == 0) return EF_OK;
    if ((rows + 31) / 32 > 65535) {
        set_error("too many rows to transpose in one launch");
        return EF_INVALID;
    }
    hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32)), dim3(256),
                       0, s, in, rows, cols, out);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status narrow_f32(const double *in, int64_t n, float *out, hipStream_t s) {
    if (n == 0) return EF_OK;
    hipLaunchKernelGGL(narrow_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, in, n, out);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

}  // namespace ef
