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
// One query row per thread, its J tiles in order (fused.rs:694-1019).
// out[x][y] is the row's running sum a; special[x][y] marks values that
// were infinite or NaN.
__global__ __launch_bounds__(64) void pipeline_rows_kernel(PipeArgs p) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= p.nx) return;
    const int64_t ny = p.ny, nj = p.nj, ne = p.ne;
    double *a = p.out + x * ny;
    uint8_t *special_value = p.special + x * ny;
    for (int64_t y = 0; y < ny; y++) {
        a[y] = 0.0;
        special_value[y] = 0;
    }
    const bool mapped = p.n_weight > 0;
    const bool shifted = !mapped && p.shifted;
    const bool unnormalized = p.normalizer == EF_NORM_NONE;
    double m = shifted ? -INFINITY : 0.0, z = 0.0, high = -INFINITY, light = INFINITY;
    double row_largest = 0.0, row_smallest = INFINITY, low = INFINITY;
    bool special_weight = false, over = false;
    uint64_t terms = 0;
    const int64_t lo_x = p.lo[x], hi_x = p.hi[x];
    double sc[JT];
    double qv[8], kv[8];
    if (lo_x < hi_x) {
        for (int64_t j0 = lo_x - lo_x % JT; j0 < hi_x; j0 += JT) {
            const int64_t width = JT < nj - j0 ? JT : nj - j0;
            const int64_t lo = (lo_x < j0 ? j0 : lo_x) - j0;
            const int64_t hi = (hi_x > j0 + width ? j0 + width : hi_x) - j0;
            if (lo >= hi) continue;
            const int n = (int)(hi - lo);
            // The producer's scores (fused.rs:771-811).
            for (int i = 0; i < n; i++) {
                const int64_t j = j0 + lo + i;
                double s = 0.0;
                if (p.n_score == 0) {
                    for (int64_t e = 0; e < ne; e++) s += p.q[x * ne + e] * p.kt[e * nj + j];
                } else {
                    for (int64_t e = 0; e < ne; e++) {
                        for (int c = 0; c < p.n_q_cols; c++) qv[c] = p.q_cols[c][x * ne + e];
                        for (int c = 0; c < p.n_k_cols; c++) kv[c] = p.k_cols[c][e * nj + j];
                        const double t = eval(p.score, p.n_score, qv, kv, 0.0);
                        s = e == 0 ? t : s + t;
                    }
                }
                sc[i] = s;
            }
            terms += (uint64_t)n;
            for (int i = 0; i < n; i++) {
                row_largest = fmax(row_largest, p.largest[j0 + lo + i]);
                row_smallest = fmin(row_smallest, p.smallest[j0 + lo + i]);
            }
            if (mapped) {
                for (int i = 0; i < n; i++) sc[i] = eval(p.weight, p.n_weight, nullptr, nullptr, sc[i]);
                for (int i = 0; i < n; i++)
                    if (!isfinite(sc[i]) || !p.finite[j0 + lo + i]) over = true;
                for (int i = 0; i < n; i++) {
                    const double w = fabs(sc[i]);
                    high = fmax(high, w);
                    if (w != 0.0) light = fmin(light, w);
                }
                double sum = -0.0;   // Rust's f64 Sum starts from -0.0
                for (int i = 0; i < n; i++) sum += sc[i];
                z += sum;
            } else {
                for (int i = 0; i < n; i++) {
                    sc[i] = scale_apply(p, sc[i]);
                    if (isfinite(sc[i])) {
                        high = fmax(high, sc[i]);
                        light = fmin(light, sc[i]);
                    }
                }
                for (int i = 0; i < n; i++) {
                    const int64_t j = j0 + lo + i;
                    if (!p.finite[j] && p.has_inf[j] && isfinite(sc[i])) low = fmin(low, sc[i]);
                }
                double shift = 0.0;
                if (shifted) {
                    shift = m;
                    for (int i = 0; i < n; i++)
                        if (sc[i] > shift || isnan(sc[i])) shift = sc[i];
                }
                if (shift == -INFINITY) {
                    for (int i = 0; i < n; i++) sc[i] = 0.0;
                } else {
                    const double rescale = exp(m - shift);
                    double sum = 0.0;
                    bool special_weights = false;
                    for (int i = 0; i < n; i++) {
                        sc[i] = exp(sc[i] - shift);
                        sum += sc[i];
                        if (!isfinite(sc[i])) special_weights = true;
                    }
                    if (shifted && rescale != 1.0) {
                        if (overflows(z, a, special_weight, special_value, ny)) over = true;
                        z *= rescale;
                        for (int64_t y = 0; y < ny; y++) a[y] *= rescale;
                    }
                    special_weight |= special_weights;
                    bool special_values = false;
                    for (int i = 0; i < n; i++) {
                        const int64_t j = j0 + lo + i;
                        if (p.finite[j]) continue;
                        special_values = true;
                        for (int64_t y = 0; y < ny; y++)
                            if (!isfinite(p.v[j * ny + y])) special_value[y] = 1;
                    }
                    if (unnormalized && (special_weights || special_values)) over = true;
                    z += sum;
                    m = shift;
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
    // Normalize (fused.rs:941-964).
    if (unnormalized) {
        if (shifted && m == -INFINITY)
            for (int64_t y = 0; y < ny; y++) a[y] = NAN;
    } else {
        if (overflows(z, a, special_weight, special_value, ny)) over = true;
        for (int64_t y = 0; y < ny; y++) a[y] /= z;
        if (z != 0.0 && overflows(z, a, special_weight, special_value, ny)) over = true;
    }
    // Underflow and overflow guards (fused.rs:980-1013).
    double heaviest = mapped ? high : (shifted ? exp(high - m) : exp(high));
    if (!(special_weight || high == -INFINITY || heaviest == 0.0)) {
        const double lightest = mapped ? light : (shifted ? exp(light - m) : exp(light));
        if (!bounds_hold(heaviest, lightest, row_largest, row_smallest, terms, !unnormalized, z, a, special_value, ny))
            over = true;
        if (low != INFINITY) {
            const double lw = shifted ? exp(low - m) : exp(low);
            const double normal = 2.0 * DBL_MIN;
            if (!(lw >= normal && (unnormalized || lw / fabs(z) >= normal))) over = true;
        }
    }
    if (over) atomicOr(p.overflowed, 1);
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

ef_status pipeline_rows(const PipeArgs &p, hipStream_t s) {
    if (p.nx == 0) return EF_OK;
    hipLaunchKernelGGL(pipeline_rows_kernel, dim3(blocks_for(p.nx, 64)), dim3(64), 0, s, p);
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
