// Device kernels of the discrimination algorithm (crates/einfold-kernels/
// src/discriminate.rs:240-650): key columns as order-preserving words for
// radix ranking, join filters, the counting sorts' keys and tie-breaks,
// segment starts, and Gustavson's product.
//
// Gustavson's product keeps the reference's summation order exactly: one
// wave owns an x rank at a time and walks its rows of a in their sorted
// order; for each, the lanes split the z bucket of b by runs of equal y (b
// is sorted by (z, y, factor bits), so a run is contiguous) and each run's
// lane adds its products in order.  A wave-scope fence separates rows of
// a, so every y's sum is the reference's left-to-right sum.  Products and
// sums are rounded separately (no FMA contraction), as Rust's are.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

#include "common.h"
#include "discriminate.h"
#include "ops.h"

namespace ef {

namespace {

// ---- key words --------------------------------------------------------

// An order-preserving word of one key (discriminate.rs:212-235, the row
// format's order for one column): signed integers with the sign bit
// flipped after widening, floats by their total order (negative: all bits
// flipped; else the sign bit set).  NULL rows get 0; the caller orders
// NULLs first separately.
template <class T>
__global__ __launch_bounds__(256) void order_words_kernel(const T *__restrict__ raw, const uint8_t *__restrict__ valid,
                                                          int64_t n, uint64_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (valid && !valid[i]) {
        out[i] = 0;
        return;
    }
    const T v = raw[i];
    uint64_t w;
    if constexpr (std::is_same<T, double>::value) {
        uint64_t b = __double_as_longlong(v);
        w = (b >> 63) ? ~b : (b | (1ULL << 63));
    } else if constexpr (std::is_same<T, float>::value) {
        uint32_t b = __float_as_uint(v);
        w = (b >> 31) ? (uint64_t)(~b) : (uint64_t)(b | 0x80000000u);
    } else if constexpr (std::is_signed<T>::value) {
        w = (uint64_t)(int64_t)v ^ (1ULL << 63);
    } else {
        w = (uint64_t)v;
    }
    out[i] = w;
}

// keys.rs:61-76: -0.0 as 0.0, where the host compares them equal.
template <class T> __global__ __launch_bounds__(256) void zero_sign_kernel(T *raw, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && raw[i] == (T)0) raw[i] = (T)0;
}

// A nullable column's rank with NULL below every value: 0 for NULL, else
// 2 r + 1 (re-ranked afterwards to be dense).
__global__ __launch_bounds__(256) void null_code_kernel(const uint32_t *rank, const uint8_t *valid, int64_t n,
                                                        uint64_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = valid[i] ? 2ULL * rank[i] + 1ULL : 0ULL;
}

// Two dense ranks as one word, the first more significant.
__global__ __launch_bounds__(256) void combine_kernel(const uint32_t *hi, const uint32_t *lo, int64_t n,
                                                      uint64_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ((uint64_t)hi[i] << 32) | lo[i];
}

__global__ __launch_bounds__(256) void and_valid_kernel(uint8_t *join, const uint8_t *valid, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !valid[i]) join[i] = 0;
}

__global__ __launch_bounds__(256) void flags_u8_kernel(const uint8_t *in, int64_t n, uint32_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] ? 1u : 0u;
}

__global__ __launch_bounds__(256) void select_rows_kernel(const uint8_t *keep, const uint32_t *excl, int64_t n,
                                                          uint32_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && keep[i]) out[excl[i]] = (uint32_t)i;
}

// ---- sort keys ----------------------------------------------------------

__global__ __launch_bounds__(256) void pair_word_kernel(const uint32_t *rows, int64_t n, const uint32_t *hi,
                                                        const uint32_t *lo, uint64_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = rows[i];
    out[i] = ((uint64_t)hi[r] << 32) | lo[r];
}

// The tie-break key (discriminate.rs:313-330): the factor's bits, NULL
// first (two stable passes: bits with NULL as 0, then validity).
__global__ __launch_bounds__(256) void bits_word_kernel(const uint32_t *rows, int64_t n, const uint8_t *valid,
                                                        const double *value, uint64_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = rows[i];
    out[i] = valid[r] ? (uint64_t)__double_as_longlong(value[r]) : 0ULL;
}

__global__ __launch_bounds__(256) void valid_word_kernel(const uint32_t *rows, int64_t n, const uint8_t *valid,
                                                         uint64_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = valid[rows[i]] ? 1ULL : 0ULL;
}

// Whether any sorted word repeats its predecessor.  One atomic per wave at
// most: atomics on one address serialize, and repeats can be most rows.
__global__ __launch_bounds__(256) void repeated_kernel(const uint64_t *w, int64_t n, int *any) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool rep = i > 0 && i < n && w[i] == w[i - 1];
    if (__any(rep) && (threadIdx.x & 63) == 0) atomicOr(any, 1);
}

// start[k] = the first sorted index whose key (w >> 32) is at least k, for
// k = 0..nk; thread i covers the keys between its predecessor's and its own.
__global__ __launch_bounds__(256) void segment_starts_kernel(const uint64_t *w, int64_t n, int64_t nk,
                                                             uint32_t *start) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int64_t lo = i == 0 ? 0 : (int64_t)(w[i - 1] >> 32) + 1;
    const int64_t hi = i == n ? nk : (int64_t)(w[i] >> 32);
    for (int64_t k = lo; k <= hi; k++) start[k] = (uint32_t)i;
}

// Sorted rows' attributes gathered into sort order.
__global__ __launch_bounds__(256) void gather_sorted_kernel(const uint32_t *rows, const uint64_t *w, int64_t n,
                                                            const uint8_t *valid, const double *value,
                                                            uint32_t *low, uint8_t *valid_s, double *value_s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = rows[i];
    low[i] = (uint32_t)w[i];
    valid_s[i] = valid[r];
    value_s[i] = value[r];
}

// Each x's bound on its groups: min(ny, its joined pairs); and the pairs
// in all (discriminate.rs:530-545).
__global__ __launch_bounds__(256) void bounds_kernel(const uint32_t *a_start, const uint32_t *a_z,
                                                     const uint32_t *b_start, int64_t nx, int64_t ny,
                                                     uint64_t *bound, unsigned long long *pairs) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t p = 0;
    if (x < nx) {
        for (uint32_t i = a_start[x]; i < a_start[x + 1]; i++) {
            const uint32_t z = a_z[i];
            p += b_start[z + 1] - b_start[z];
        }
        bound[x] = p < (uint64_t)ny ? p : (uint64_t)ny;
    }
    for (int o = 32; o >= 1; o >>= 1) p += __shfl_xor(p, o);
    if ((threadIdx.x & 63) == 0 && p) atomicAdd(pairs, (unsigned long long)p);
}

// ---- Gustavson's product ------------------------------------------------

constexpr int CAP = 1024;   // touched y's a wave sorts in LDS; more: scan

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

#pragma clang fp contract(off)
__global__ __launch_bounds__(256) void gustavson_kernel(GustavsonArgs g) {
    __shared__ uint32_t touched[4][CAP];
    __shared__ uint32_t count[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t slot = (int64_t)blockIdx.x * 4 + wv;
    double *acc = g.acc + slot * g.ny;
    uint8_t *state = g.state + slot * g.ny;       // bit 0 reached, bit 1 valued
    const uint64_t below = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    for (;;) {
        int64_t x = 0;
        if (lane == 0) x = (int64_t)atomicAdd(g.next, 1ULL);
        x = __shfl(x, 0);
        if (x >= g.nx) break;
        if (lane == 0) count[wv] = 0;
        wave_sync_lds();
        // The rows of a, up to 64 at a time, one per lane with its bucket of
        // b, broadcast row by row; a run head's row of b read with its run
        // test.  The same products in the same order.
        const uint32_t ia = g.a_start[x], ie = g.a_start[x + 1];
        for (uint32_t r0 = ia; r0 < ie; r0 += 64) {
            const uint32_t nr = ie - r0 < 64 ? ie - r0 : 64;
            int rav = 0;
            double rp = 0.0;
            uint32_t rbs = 0, rbe = 0;
            if ((uint32_t)lane < nr) {
                const uint32_t z = g.a_z[r0 + lane];
                rav = g.a_valid[r0 + lane] != 0;
                rp = g.a_value[r0 + lane];
                rbs = g.b_start[z];
                rbe = g.b_start[z + 1];
            }
            for (uint32_t r = 0; r < nr; r++) {
                const bool av = __shfl(rav, (int)r) != 0;
                const double p = __shfl(rp, (int)r);
                const uint32_t bs = (uint32_t)__shfl((int)rbs, (int)r), be = (uint32_t)__shfl((int)rbe, (int)r);
                for (uint32_t j = bs + lane; j < be; j += 64) {
                    const uint32_t y = g.b_y[j];
                    const uint32_t yp = g.b_y[j != bs ? j - 1 : j];
                    const uint8_t bv = g.b_valid[j];
                    const double bx = g.b_value[j];
                    if (j != bs && yp == y) continue;   // not the head of its run
                    uint8_t st = state[y];
                    if (!(st & 1)) {
                        st |= 1;
                        const uint32_t t = atomicAdd(&count[wv], 1u);
                        if (t < CAP) touched[wv][t] = y;
                    }
                    double a = acc[y];
                    if (av && bv) {
                        a = a + p * bx;
                        st |= 2;
                    }
                    for (uint32_t k = j + 1; k < be && g.b_y[k] == y; k++) {
                        if (av && g.b_valid[k]) {
                            a = a + p * g.b_value[k];
                            st |= 2;
                        }
                    }
                    acc[y] = a;
                    state[y] = st;
                }
                // This row's sums land before the next row's lanes read them.
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
            }
        }
        wave_sync_lds();
        const uint32_t nt = count[wv];
        const uint64_t base = g.bound_off[x];
        if (nt <= (uint32_t)CAP) {
            // Bitonic sort of the touched y's (distinct) in LDS.
            uint32_t pw = 1;
            while (pw < nt) pw <<= 1;
            for (uint32_t t = nt + lane; t < pw; t += 64) touched[wv][t] = 0xffffffffu;
            wave_sync_lds();
            for (uint32_t k = 2; k <= pw; k <<= 1) {
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t t = lane; t < pw; t += 64) {
                        const uint32_t l = t ^ j;
                        if (l > t) {
                            const uint32_t u = touched[wv][t], v = touched[wv][l];
                            const bool up = (t & k) == 0;
                            if (up ? u > v : u < v) {
                                touched[wv][t] = v;
                                touched[wv][l] = u;
                            }
                        }
                    }
                    wave_sync_lds();
                }
            }
            for (uint32_t t = lane; t < nt; t += 64) {
                const uint32_t y = touched[wv][t];
                g.g_y[base + t] = y;
                g.g_sum[base + t] = acc[y];
                g.g_valued[base + t] = (state[y] >> 1) & 1;
                acc[y] = 0.0;
                state[y] = 0;
            }
        } else {
            // Many groups: their y's in order from a scan of the states.
            uint32_t placed = 0;
            for (int64_t y0 = 0; y0 < g.ny; y0 += 64) {
                const int64_t y = y0 + lane;
                const uint8_t st = y < g.ny ? state[y] : 0;
                const uint64_t m = __ballot(st & 1);
                if (st & 1) {
                    const uint64_t o = base + placed + __popcll(m & below);
                    g.g_y[o] = (uint32_t)y;
                    g.g_sum[o] = acc[y];
                    g.g_valued[o] = (st >> 1) & 1;
                    acc[y] = 0.0;
                    state[y] = 0;
                }
                placed += (uint32_t)__popcll(m);
            }
        }
        if (lane == 0) g.g_count[x] = nt;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
}

// The same product with few output columns (ny <= YCAP): each wave's
// accumulator and states live in LDS, rows are ordered by a wave-level
// sync instead of a memory fence, and an x's groups come out by a scan of
// the states in y order (the order of the sorted list above).  Waves take
// x's XCHUNK at a time.  The sums are the same, in the same order.
constexpr int XCHUNK = 8;

template <int YCAP>
__global__ __launch_bounds__(256) void gustavson_lds_kernel(GustavsonArgs g) {
    __shared__ double acc_s[4][YCAP];
    __shared__ uint8_t state_s[4][YCAP];   // bit 0 reached, bit 1 valued
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double *acc = acc_s[wv];
    uint8_t *state = state_s[wv];
    const uint64_t below = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    for (int64_t y = lane; y < g.ny; y += 64) {
        acc[y] = 0.0;
        state[y] = 0;
    }
    wave_sync_lds();
    for (;;) {
        int64_t x0 = 0;
        if (lane == 0) x0 = (int64_t)atomicAdd(g.next, (unsigned long long)XCHUNK);
        x0 = __shfl(x0, 0);
        if (x0 >= g.nx) break;
        const int64_t x1 = x0 + XCHUNK < g.nx ? x0 + XCHUNK : g.nx;
        for (int64_t x = x0; x < x1; x++) {
            const uint32_t ia = g.a_start[x], ie = g.a_start[x + 1];
            // The x's rows of a, up to 64 at a time, one per lane with its
            // bucket of b: read together instead of one chain per row.
            for (uint32_t i0 = ia; i0 < ie; i0 += 64) {
                const uint32_t nr = ie - i0 < 64 ? ie - i0 : 64;
                int rav = 0;
                double rp = 0.0;
                uint32_t rbs = 0, rbe = 0;
                if ((uint32_t)lane < nr) {
                    const uint32_t z = g.a_z[i0 + lane];
                    rav = g.a_valid[i0 + lane] != 0;
                    rp = g.a_value[i0 + lane];
                    rbs = g.b_start[z];
                    rbe = g.b_start[z + 1];
                }
                for (uint32_t r = 0; r < nr; r++) {
                    const bool av = __shfl(rav, (int)r) != 0;
                    const double p = __shfl(rp, (int)r);
                    const uint32_t bs = (uint32_t)__shfl((int)rbs, (int)r), be = (uint32_t)__shfl((int)rbe, (int)r);
                    for (uint32_t j = bs + lane; j < be; j += 64) {
                        // The head's row of b read with the test of its run
                        // (the row before it from a valid index).
                        const uint32_t y = g.b_y[j];
                        const uint32_t yp = g.b_y[j != bs ? j - 1 : j];
                        const uint8_t bv = g.b_valid[j];
                        const double bx = g.b_value[j];
                        if (j != bs && yp == y) continue;   // not the head of its run
                        uint8_t st = state[y] | 1;
                        double a = acc[y];
                        if (av && bv) {
                            a = a + p * bx;
                            st |= 2;
                        }
                        for (uint32_t k = j + 1; k < be && g.b_y[k] == y; k++) {
                            if (av && g.b_valid[k]) {
                                a = a + p * g.b_value[k];
                                st |= 2;
                            }
                        }
                        acc[y] = a;
                        state[y] = st;
                    }
                    // This row's sums land before the next row's lanes read them.
                    wave_sync_lds();
                }
            }
            const uint64_t base = g.bound_off[x];
            uint32_t placed = 0;
            for (int64_t y0 = 0; y0 < g.ny; y0 += 64) {
                const int64_t y = y0 + lane;
                const uint8_t st = y < g.ny ? state[y] : 0;
                const uint64_t m = __ballot(st & 1);
                if (st & 1) {
                    const uint64_t o = base + placed + __popcll(m & below);
                    g.g_y[o] = (uint32_t)y;
                    g.g_sum[o] = acc[y];
                    g.g_valued[o] = (st >> 1) & 1;
                    acc[y] = 0.0;
                    state[y] = 0;
                }
                placed += (uint32_t)__popcll(m);
            }
            if (lane == 0) g.g_count[x] = placed;
            wave_sync_lds();
        }
    }
}
// ---- Gustavson's product with heavy join ranks -------------------------
//
// Zipf-distributed join keys put most joined pairs in a few z buckets (the
// hottest j of 100k rows per side holds two thirds of them).  Walking such
// a bucket once per row of a, as above, is a read-modify-write of the
// accumulator per pair.  Here every (x, y) cell has its own accumulator and
// state in HBM, and the product runs in phases in z order: the light z's
// below the first heavy one (a wave per x, as above), the first heavy z (a
// thread per (x, run of equal y in its bucket): the cell's sum in a
// register over the x's rows of a at that z, each times the run's rows of
// b, both in order), the light z's up to the next heavy one, and so on.
// Every cell therefore adds the same products in the same order as the
// wave-per-x loop: z ascending, then a's rows, then b's rows.  The groups
// come out by a scan of each x's states in y order.

// split[2 k][x], split[2 k + 1][x]: the first row of x's a rows whose z is
// at least heavy[k], and above it (a is sorted by (x, z)).
__global__ __launch_bounds__(256) void heavy_split_kernel(const uint32_t *a_start, const uint32_t *a_z, int64_t nx,
                                                          const uint32_t *heavy, int n_heavy, uint32_t *split) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx) return;
    const uint32_t lo0 = a_start[x], hi0 = a_start[x + 1];
    for (int k = 0; k < n_heavy; k++) {
        const uint32_t h = heavy[k];
        uint32_t lo = lo0, hi = hi0;
        while (lo < hi) {   // first z >= h
            const uint32_t mid = (lo + hi) >> 1;
            if (a_z[mid] < h) lo = mid + 1;
            else hi = mid;
        }
        uint32_t e = lo, eh = hi0;
        while (e < eh) {   // first z > h
            const uint32_t mid = (e + eh) >> 1;
            if (a_z[mid] <= h) e = mid + 1;
            else eh = mid;
        }
        split[(2 * (int64_t)k) * nx + x] = lo;
        split[(2 * (int64_t)k + 1) * nx + x] = e;
    }
}

// The light z's of a phase: a wave per x over its a rows [lo[x], hi[x])
// (null: the x's first or last row), into the x's own cells.
__global__ __launch_bounds__(256) void gustavson_light_kernel(GustavsonArgs g, const uint32_t *lo_idx,
                                                              const uint32_t *hi_idx) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t x = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); x < g.nx; x += waves) {
        double *acc = g.acc + x * g.ny;
        uint8_t *state = g.state + x * g.ny;
        const uint32_t i0 = lo_idx ? lo_idx[x] : g.a_start[x];
        const uint32_t i1 = hi_idx ? hi_idx[x] : g.a_start[x + 1];
        // The rows of a, up to 64 at a time, one per lane with its bucket of
        // b, broadcast row by row; a run head's row of b read with its run
        // test.  The same products in the same order.
        for (uint32_t r0 = i0; r0 < i1; r0 += 64) {
            const uint32_t nr = i1 - r0 < 64 ? i1 - r0 : 64;
            int rav = 0;
            double rp = 0.0;
            uint32_t rbs = 0, rbe = 0;
            if ((uint32_t)lane < nr) {
                const uint32_t z = g.a_z[r0 + lane];
                rav = g.a_valid[r0 + lane] != 0;
                rp = g.a_value[r0 + lane];
                rbs = g.b_start[z];
                rbe = g.b_start[z + 1];
            }
            for (uint32_t r = 0; r < nr; r++) {
                const bool av = __shfl(rav, (int)r) != 0;
                const double p = __shfl(rp, (int)r);
                const uint32_t bs = (uint32_t)__shfl((int)rbs, (int)r), be = (uint32_t)__shfl((int)rbe, (int)r);
                for (uint32_t j = bs + lane; j < be; j += 64) {
                    const uint32_t y = g.b_y[j];
                    const uint32_t yp = g.b_y[j != bs ? j - 1 : j];
                    const uint8_t bv = g.b_valid[j];
                    const double bx = g.b_value[j];
                    if (j != bs && yp == y) continue;   // not the head of its run
                    uint8_t st = state[y] | 1;
                    double a = acc[y];
                    if (av && bv) {
                        a = a + p * bx;
                        st |= 2;
                    }
                    for (uint32_t k = j + 1; k < be && g.b_y[k] == y; k++) {
                        if (av && g.b_valid[k]) {
                            a = a + p * g.b_value[k];
                            st |= 2;
                        }
                    }
                    acc[y] = a;
                    state[y] = st;
                }
                // This row's sums land before the next row's lanes read them.
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
}

// The first row of each run of equal y in heavy z's b bucket, in order,
// and their count (one workgroup of 1024 threads per heavy z).
__global__ __launch_bounds__(1024) void heavy_runs_kernel(const uint32_t *b_start, const uint32_t *b_y,
                                                          const uint32_t *heavy, const uint64_t *run_off,
                                                          uint32_t *run_start, uint32_t *run_count) {
    __shared__ uint32_t wave_n[16];
    __shared__ uint32_t base;
    const uint32_t h = heavy[blockIdx.x];
    const uint32_t bs = b_start[h], be = b_start[h + 1];
    uint32_t *out = run_start + run_off[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) base = 0;
    __syncthreads();
    for (uint32_t j0 = bs; j0 < be; j0 += 1024) {
        const uint32_t j = j0 + tid;
        const bool head = j < be && (j == bs || b_y[j - 1] != b_y[j]);
        const uint64_t m = __ballot(head);
        if (lane == 0) wave_n[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = base;
        for (int w = 0; w < wv; w++) before += wave_n[w];
        if (head) out[before + __popcll(m & (lane == 0 ? 0ULL : (~0ULL >> (64 - lane))))] = j;
        __syncthreads();
        if (tid == 0) {
            uint32_t t = 0;
            for (int w = 0; w < 16; w++) t += wave_n[w];
            base += t;
        }
        __syncthreads();
    }
    if (tid == 0) run_count[blockIdx.x] = base;
}

// One heavy z: a thread per run of equal y in its bucket (runs [r0, r0 +
// 256) of block r0 / 256 in each group of HX x's: blockIdx.x = x group *
// per_x + block) sums the run against each x's a rows at z.  Every cell's
// word is loaded before any is stored, so the loads of the HX cells are in
// flight together.
constexpr int HX = 16;

__global__ __launch_bounds__(256) void gustavson_heavy_kernel(GustavsonArgs g, uint32_t h, const uint32_t *lo_idx,
                                                              const uint32_t *hi_idx, const uint32_t *runs,
                                                              const uint32_t *n_runs, int64_t per_x) {
    const int64_t x0 = (int64_t)blockIdx.x / per_x * HX;
    const int64_t r = ((int64_t)blockIdx.x % per_x) * 256 + threadIdx.x;
    const uint32_t nr = *n_runs;
    if (r >= nr) return;
    const uint32_t be = g.b_start[h + 1];
    const uint32_t j = runs[r], e = r + 1 < nr ? runs[r + 1] : be;
    const uint32_t y = g.b_y[j];
    uint32_t i0[HX], i1[HX];
    double a[HX];
    uint8_t st[HX];
#pragma unroll
    for (int xi = 0; xi < HX; xi++) {
        const int64_t x = x0 + xi < g.nx ? x0 + xi : g.nx - 1;
        i0[xi] = lo_idx[x];
        i1[xi] = x0 + xi < g.nx ? hi_idx[x] : i0[xi];
    }
#pragma unroll
    for (int xi = 0; xi < HX; xi++) {
        if (i0[xi] == i1[xi]) continue;
        a[xi] = g.acc[(x0 + xi) * g.ny + y];
        st[xi] = g.state[(x0 + xi) * g.ny + y];
    }
    // A short run's factors in registers (NULL as a zero flag), read once.
    constexpr int RUN = 8;
    const bool short_run = e - j <= (uint32_t)RUN;
    double bv[RUN];
    bool bok[RUN];
#pragma unroll
    for (int q = 0; q < RUN; q++) {
        const uint32_t k = j + q < e ? j + q : j;
        bv[q] = g.b_value[k];
        bok[q] = j + q < e && g.b_valid[k] != 0;
    }
#pragma unroll
    for (int xi = 0; xi < HX; xi++) {
        if (i0[xi] == i1[xi]) continue;
        double s = a[xi];
        uint8_t t = st[xi] | 1;
        for (uint32_t i = i0[xi]; i < i1[xi]; i++) {
            const bool av = g.a_valid[i] != 0;
            const double p = g.a_value[i];
            if (short_run) {
#pragma unroll
                for (int q = 0; q < RUN; q++) {
                    if (av && bok[q]) {
                        s = s + p * bv[q];
                        t |= 2;
                    }
                }
            } else {
                for (uint32_t k = j; k < e; k++) {
                    if (av && g.b_valid[k]) {
                        s = s + p * g.b_value[k];
                        t |= 2;
                    }
                }
            }
        }
        g.acc[(x0 + xi) * g.ny + y] = s;
        g.state[(x0 + xi) * g.ny + y] = t;
    }
}

// Each x's groups from its cells' states, in y order, into its staged
// span (bound_off[x]), as the wave-per-x loop leaves them.
__global__ __launch_bounds__(256) void collect_groups_kernel(GustavsonArgs g) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t x = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); x < g.nx; x += waves) {
        const double *acc = g.acc + x * g.ny;
        const uint8_t *state = g.state + x * g.ny;
        const uint64_t base = g.bound_off[x];
        uint32_t placed = 0;
        for (int64_t y0 = 0; y0 < g.ny; y0 += 64) {
            const int64_t y = y0 + lane;
            const uint8_t st = y < g.ny ? state[y] : 0;
            const uint64_t m = __ballot(st & 1);
            if (st & 1) {
                const uint64_t o = base + placed + __popcll(m & below);
                g.g_y[o] = (uint32_t)y;
                g.g_sum[o] = acc[y];
                g.g_valued[o] = (st >> 1) & 1;
            }
            placed += (uint32_t)__popcll(m);
        }
        if (lane == 0) g.g_count[x] = placed;
    }
}
#pragma clang fp contract(on)

// ---- output ----------------------------------------------------------------

template <int W> struct Word;
template <> struct Word<1> { typedef uint8_t T; };
template <> struct Word<2> { typedef uint16_t T; };
template <> struct Word<4> { typedef uint32_t T; };
template <> struct Word<8> { typedef uint64_t T; };

__device__ __forceinline__ void copy_key(const GroupKeys &k, int i, int64_t src, int64_t dst) {
    switch (k.width[i]) {
    case 1: static_cast<uint8_t *>(k.dst[i])[dst] = static_cast<const uint8_t *>(k.src[i])[src]; break;
    case 2: static_cast<uint16_t *>(k.dst[i])[dst] = static_cast<const uint16_t *>(k.src[i])[src]; break;
    case 4: static_cast<uint32_t *>(k.dst[i])[dst] = static_cast<const uint32_t *>(k.src[i])[src]; break;
    default: static_cast<uint64_t *>(k.dst[i])[dst] = static_cast<const uint64_t *>(k.src[i])[src]; break;
    }
    k.dst_valid[i][dst] = k.src_valid[i][src];
}

// b's output keys by y rank: key i's value and validity at the first row
// of each y rank, rank r at index r (one thread per rank and key).
__global__ __launch_bounds__(256) void y_keys_kernel(int64_t ny, const uint32_t *y_first, GroupKeys src,
                                                     GroupKeys by_rank) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ny) return;
    for (int i = 0; i < src.n; i++) {
        if (src.from_a[i]) continue;
        GroupKeys k = src;
        k.dst[i] = const_cast<void *>(by_rank.src[i]);
        k.dst_valid[i] = const_cast<uint8_t *>(by_rank.src_valid[i]);
        copy_key(k, i, y_first[r], r);
    }
}

// The output's rows, one wave per x: each group's keys taken from the
// first row of its x rank (a's outputs) or, already gathered by rank, its y
// rank (b's), and its sum or NULL (discriminate.rs:586-610).
__global__ __launch_bounds__(256) void emit_groups_kernel(int64_t nx, const uint64_t *bound_off,
                                                          const uint64_t *out_off, const uint32_t *g_count,
                                                          const uint32_t *g_y, const double *g_sum,
                                                          const uint8_t *g_valued, const uint32_t *x_first,
                                                          GroupKeys keys, double *values, uint8_t *valid) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t x = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); x < nx; x += waves) {
        const uint32_t n = g_count[x];
        const uint64_t src = bound_off[x], dst = out_off[x];
        const uint32_t xr = x_first[x];
        for (uint32_t t = lane; t < n; t += 64) {
            const uint32_t y = g_y[src + t];
            const int64_t o = (int64_t)(dst + t);
            for (int i = 0; i < keys.n; i++) copy_key(keys, i, keys.from_a[i] ? xr : y, o);
            values[o] = g_sum[src + t];
            valid[o] = g_valued[src + t];
        }
    }
}

// A domain row's value: its group's sum, found by its ranks, where its
// keys can match and the group has a value; else the default or NULL
// (discriminate.rs:611-648).
__global__ __launch_bounds__(256) void domain_lookup_kernel(int64_t n, const uint8_t *matches, const uint32_t *x_rank,
                                                            const uint32_t *y_rank, int64_t nx,
                                                            const uint64_t *bound_off, const uint32_t *g_count,
                                                            const uint32_t *g_y, const double *g_sum,
                                                            const uint8_t *g_valued, int has_missing, double missing,
                                                            double *values, uint8_t *valid) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    bool got = false;
    double v = 0.0;
    const uint32_t xr = x_rank[r], yr = y_rank[r];
    if (matches[r] && (int64_t)xr < nx) {
        uint64_t lo = bound_off[xr], hi = lo + g_count[xr];
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (g_y[mid] < yr) lo = mid + 1;
            else hi = mid;
        }
        if (lo < bound_off[xr] + g_count[xr] && g_y[lo] == yr && g_valued[lo]) {
            got = true;
            v = g_sum[lo];
        }
    }
    values[r] = got ? v : (has_missing ? missing : 0.0);
    valid[r] = (got || has_missing) ? 1 : 0;
}

// ---- u64 exclusive scan ------------------------------------------------------

__global__ __launch_bounds__(256) void scan64_block_kernel(const uint64_t *in, uint64_t *out, uint64_t *sums,
                                                           int64_t n) {
    __shared__ uint64_t wave_tot[4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * 1024 + tid * 4;
    uint64_t v[4], t = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i] = base + i < n ? in[base + i] : 0;
        t += v[i];
    }
    uint64_t inc = t;
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t o = __shfl_up(inc, off);
        if (lane >= off) inc += o;
    }
    if (lane == 63) wave_tot[wv] = inc;
    __syncthreads();
    uint64_t prefix = 0;
    for (int k = 0; k < wv; k++) prefix += wave_tot[k];
    uint64_t run = prefix + inc - t;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (sums && tid == 255) sums[blockIdx.x] = prefix + inc;
}

__global__ __launch_bounds__(256) void add64_kernel(uint64_t *out, const uint64_t *sums, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / 1024];
}

__global__ __launch_bounds__(256) void u32_to_u64_kernel(const uint32_t *in, int64_t n, uint64_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

size_t scan64_scratch_bytes(size_t n) {
    size_t bytes = 256;
    while (n > 1024) {
        n = (n + 1023) / 1024;
        bytes += up256(n * 8);
    }
    return bytes;
}

ef_status scan_u64(const uint64_t *in, uint64_t *out, size_t n, void *tmp, hipStream_t s) {
    if (n == 0) return EF_OK;
    const int64_t blocks = (int64_t)((n + 1023) / 1024);
    uint64_t *sums = blocks > 1 ? static_cast<uint64_t *>(tmp) : nullptr;
    hipLaunchKernelGGL(scan64_block_kernel, dim3((unsigned)blocks), dim3(256), 0, s, in, out, sums, (int64_t)n);
    if (blocks > 1) {
        char *next = static_cast<char *>(tmp) + up256(blocks * 8);
        EF_TRY(scan_u64(sums, sums, (size_t)blocks, next, s));
        hipLaunchKernelGGL(add64_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, out, sums, (int64_t)n);
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status order_words(Type t, const void *raw, const uint8_t *valid, int64_t n, uint64_t *out, hipStream_t s) {
    if (n == 0) return EF_OK;
    const unsigned g = blocks_for(n, 256);
#define EF_OW(T) hipLaunchKernelGGL(order_words_kernel<T>, dim3(g), dim3(256), 0, s, (const T *)raw, valid, n, out)
    switch (t) {
    case Type::Int8: EF_OW(int8_t); break;
    case Type::UInt8: EF_OW(uint8_t); break;
    case Type::Int16: EF_OW(int16_t); break;
    case Type::UInt16: EF_OW(uint16_t); break;
    case Type::Int32: EF_OW(int32_t); break;
    case Type::UInt32: EF_OW(uint32_t); break;
    case Type::Int64: EF_OW(int64_t); break;
    case Type::UInt64: EF_OW(uint64_t); break;
    case Type::Float32: EF_OW(float); break;
    case Type::Float64: EF_OW(double); break;
    default: set_error("keys must be integers or floats"); return EF_INVALID;
    }
#undef EF_OW
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status zero_signs(Type t, void *raw, int64_t n, hipStream_t s) {
    if (n == 0) return EF_OK;
    if (t == Type::Float64)
        hipLaunchKernelGGL(zero_sign_kernel<double>, dim3(blocks_for(n, 256)), dim3(256), 0, s, (double *)raw, n);
    else if (t == Type::Float32)
        hipLaunchKernelGGL(zero_sign_kernel<float>, dim3(blocks_for(n, 256)), dim3(256), 0, s, (float *)raw, n);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

#define EF_LAUNCH(kernel, n, ...)                                                                       \
    do {                                                                                               \
        if ((n) > 0) hipLaunchKernelGGL(kernel, dim3(blocks_for((n), 256)), dim3(256), 0, s, __VA_ARGS__); \
        EF_HIP(hipGetLastError());                                                                     \
    } while (0)

ef_status null_codes(const uint32_t *rank, const uint8_t *valid, int64_t n, uint64_t *out, hipStream_t s) {
    EF_LAUNCH(null_code_kernel, n, rank, valid, n, out);
    return EF_OK;
}

ef_status combine_ranks(const uint32_t *hi, const uint32_t *lo, int64_t n, uint64_t *out, hipStream_t s) {
    EF_LAUNCH(combine_kernel, n, hi, lo, n, out);
    return EF_OK;
}

ef_status and_valid(uint8_t *join, const uint8_t *valid, int64_t n, hipStream_t s) {
    EF_LAUNCH(and_valid_kernel, n, join, valid, n);
    return EF_OK;
}

ef_status select_rows(const uint8_t *keep, int64_t n, uint32_t *flags, uint32_t *excl, uint32_t *rows,
                      void *scan_tmp, uint32_t *count, hipStream_t s) {
    *count = 0;
    if (n == 0) return EF_OK;
    EF_LAUNCH(flags_u8_kernel, n, keep, n, flags);
    EF_TRY(scan_u32(flags, excl, (size_t)n, scan_tmp, s));
    EF_LAUNCH(select_rows_kernel, n, keep, excl, n, rows);
    uint32_t last[2];
    uint8_t lk;
    EF_HIP(hipMemcpyAsync(&last[0], excl + n - 1, 4, hipMemcpyDeviceToHost, s));
    EF_HIP(hipMemcpyAsync(&lk, keep + n - 1, 1, hipMemcpyDeviceToHost, s));
    EF_HIP(hipStreamSynchronize(s));
    *count = last[0] + (lk ? 1 : 0);
    return EF_OK;
}

ef_status pair_words(const uint32_t *rows, int64_t n, const uint32_t *hi, const uint32_t *lo, uint64_t *out,
                     hipStream_t s) {
    EF_LAUNCH(pair_word_kernel, n, rows, n, hi, lo, out);
    return EF_OK;
}

ef_status tie_words(const uint32_t *rows, int64_t n, const uint8_t *valid, const double *value, uint64_t *bits,
                    uint64_t *valid_w, hipStream_t s) {
    EF_LAUNCH(bits_word_kernel, n, rows, n, valid, value, bits);
    EF_LAUNCH(valid_word_kernel, n, rows, n, valid, valid_w);
    return EF_OK;
}

ef_status any_repeated(const uint64_t *w, int64_t n, int *flag, bool *out, hipStream_t s) {
    *out = false;
    if (n < 2) return EF_OK;
    EF_HIP(hipMemsetAsync(flag, 0, 4, s));
    EF_LAUNCH(repeated_kernel, n, w, n, flag);
    int h = 0;
    EF_HIP(hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, s));
    EF_HIP(hipStreamSynchronize(s));
    *out = h != 0;
    return EF_OK;
}

ef_status segment_starts(const uint64_t *w, int64_t n, int64_t nk, uint32_t *start, hipStream_t s) {
    EF_LAUNCH(segment_starts_kernel, n + 1, w, n, nk, start);
    return EF_OK;
}

ef_status gather_sorted(const uint32_t *rows, const uint64_t *w, int64_t n, const uint8_t *valid,
                        const double *value, uint32_t *low, uint8_t *valid_s, double *value_s, hipStream_t s) {
    EF_LAUNCH(gather_sorted_kernel, n, rows, w, n, valid, value, low, valid_s, value_s);
    return EF_OK;
}

ef_status group_bounds(const uint32_t *a_start, const uint32_t *a_z, const uint32_t *b_start, int64_t nx, int64_t ny,
                       uint64_t *bound, unsigned long long *pairs, hipStream_t s) {
    EF_LAUNCH(bounds_kernel, nx, a_start, a_z, b_start, nx, ny, bound, pairs);
    return EF_OK;
}

ef_status gustavson(const GustavsonArgs &g, int blocks, hipStream_t s) {
    if (g.nx == 0) return EF_OK;
    if (g.ny <= 512) {
        const int64_t want = (g.nx + 4 * XCHUNK - 1) / (4 * XCHUNK);
        hipLaunchKernelGGL(gustavson_lds_kernel<512>, dim3((unsigned)(want < 2048 ? want : 2048)), dim3(256), 0, s, g);
    } else {
        hipLaunchKernelGGL(gustavson_kernel, dim3(blocks), dim3(256), 0, s, g);
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status gustavson_heavy(const GustavsonArgs &g, const uint32_t *heavy_host, const uint32_t *heavy_dev,
                          const uint32_t *heavy_rows, int n_heavy, uint32_t *split, uint32_t *runs,
                          const uint64_t *run_off_host, const uint64_t *run_off_dev, uint32_t *n_runs, hipStream_t s) {
    if (g.nx == 0) return EF_OK;
    const unsigned wave_blocks = (unsigned)((g.nx + 3) / 4 < 4096 ? (g.nx + 3) / 4 : 4096);
    EF_LAUNCH(heavy_split_kernel, g.nx, g.a_start, g.a_z, g.nx, heavy_dev, n_heavy, split);
    hipLaunchKernelGGL(heavy_runs_kernel, dim3((unsigned)n_heavy), dim3(1024), 0, s, g.b_start, g.b_y, heavy_dev,
                       run_off_dev, runs, n_runs);
    const int64_t groups = (g.nx + HX - 1) / HX;
    for (int k = 0; k <= n_heavy; k++) {
        const uint32_t *lo = k == 0 ? nullptr : split + (2 * (int64_t)(k - 1) + 1) * g.nx;
        const uint32_t *hi = k == n_heavy ? nullptr : split + (2 * (int64_t)k) * g.nx;
        hipLaunchKernelGGL(gustavson_light_kernel, dim3(wave_blocks), dim3(256), 0, s, g, lo, hi);
        if (k == n_heavy) break;
        const int64_t per_x = ((int64_t)heavy_rows[k] + 255) / 256;   // runs are at most the bucket's rows
        if (per_x == 0) continue;
        if (per_x * groups > 0x7fffffffLL) {
            set_error("a heavy join rank's bucket is too large for one launch");
            return EF_INVALID;
        }
        hipLaunchKernelGGL(gustavson_heavy_kernel, dim3((unsigned)(per_x * groups)), dim3(256), 0, s, g, heavy_host[k],
                           split + (2 * (int64_t)k) * g.nx, split + (2 * (int64_t)k + 1) * g.nx,
                           runs + run_off_host[k], n_runs + k, per_x);
    }
    hipLaunchKernelGGL(collect_groups_kernel, dim3(wave_blocks), dim3(256), 0, s, g);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status widen_counts(const uint32_t *in, int64_t n, uint64_t *out, hipStream_t s) {
    EF_LAUNCH(u32_to_u64_kernel, n, in, n, out);
    return EF_OK;
}

size_t emit_scratch_bytes(int64_t ny) { return 8 * ((((size_t)(ny > 0 ? ny : 1) * 8 + 255) & ~(size_t)255) + (((size_t)(ny > 0 ? ny : 1) + 255) & ~(size_t)255)); }

ef_status emit_groups(int64_t nx, int64_t ny, const uint64_t *bound_off, const uint64_t *out_off,
                      const uint32_t *g_count, const uint32_t *g_y, const double *g_sum, const uint8_t *g_valued,
                      const uint32_t *x_first, const uint32_t *y_first, const GroupKeys &keys, void *scratch,
                      double *values, uint8_t *valid, hipStream_t s) {
    if (nx == 0) return EF_OK;
    // b's keys gathered once per y rank (ny values each), so that the
    // emission reads a small table instead of scattered rows of the
    // original columns for every group.
    GroupKeys emit = keys, by_rank = keys;
    char *at = static_cast<char *>(scratch);
    bool any_b = false;
    for (int i = 0; i < keys.n; i++) {
        if (keys.from_a[i]) continue;
        any_b = true;
        by_rank.src[i] = at;
        at += ((size_t)ny * 8 + 255) & ~(size_t)255;
        by_rank.src_valid[i] = reinterpret_cast<uint8_t *>(at);
        at += ((size_t)ny + 255) & ~(size_t)255;
        emit.src[i] = by_rank.src[i];
        emit.src_valid[i] = by_rank.src_valid[i];
    }
    if (any_b && ny > 0)
        hipLaunchKernelGGL(y_keys_kernel, dim3(blocks_for(ny, 256)), dim3(256), 0, s, ny, y_first, keys, by_rank);
    const int64_t blocks = (nx + 3) / 4 < 4096 ? (nx + 3) / 4 : 4096;
    hipLaunchKernelGGL(emit_groups_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nx, bound_off, out_off, g_count,
                       g_y, g_sum, g_valued, x_first, emit, values, valid);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status domain_lookup(int64_t n, const uint8_t *matches, const uint32_t *x_rank, const uint32_t *y_rank, int64_t nx,
                        const uint64_t *bound_off, const uint32_t *g_count, const uint32_t *g_y, const double *g_sum,
                        const uint8_t *g_valued, int has_missing, double missing, double *values, uint8_t *valid,
                        hipStream_t s) {
    EF_LAUNCH(domain_lookup_kernel, n, n, matches, x_rank, y_rank, nx, bound_off, g_count, g_y, g_sum, g_valued,
              has_missing, missing, values, valid);
    return EF_OK;
}

}  // namespace ef
