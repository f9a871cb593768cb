// Device kernels of the dense positional fold's operator path
// (crates/einfold-kernels/src/dense.rs:97-136, 319-434, 468-650): keys cast
// to i64 with their extents, positions, placement with existence flags and
// duplicate detection, the domain's cells, and the output's columns.
// Everything here is integer work or a copy; claims use integer atomics,
// whose outcome (a duplicate or not) doesn't depend on order.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.h"
#include "ops.h"

namespace ef {

namespace {

// ---- keys ----------------------------------------------------------

// Keys cast to i64, with their minimum, maximum and whether any unsigned
// key is 2^63 or more.  The bounds are reduced over the wave, then over the
// workgroup's four waves in LDS, so each workgroup adds one atomic per
// bound: atomics on one address serialize (one per wave cost 78 us over
// 262k keys).
template <class T>
__global__ __launch_bounds__(256) void cast_keys_kernel(const T *__restrict__ in, int64_t n,
                                                        int64_t *__restrict__ out, KeyStats *stats) {
    __shared__ int64_t slo[4], shi[4];
    __shared__ int sbad[4];
    int64_t lo = 0, hi = -1;
    int bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T v = in[i];
        int64_t k;
        if constexpr (sizeof(T) == 8 && !std::is_signed<T>::value) {
            // Unsigned keys of 2^63 or more don't fit i64 (dense.rs:101-108).
            if ((uint64_t)v >> 63) { bad = 1; k = 0; }
            else k = (int64_t)v;
        } else {
            k = (int64_t)v;
        }
        out[i] = k;
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        const int64_t l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
        bad |= __shfl_xor(bad, o);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        slo[w] = lo;
        shi[w] = hi;
        sbad[w] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 4; i++) {
            lo = slo[i] < lo ? slo[i] : lo;
            hi = shi[i] > hi ? shi[i] : hi;
            bad |= sbad[i];
        }
        // A bound only moves one way, so a read of it (perhaps older, never
        // tighter) that already covers this block's makes the atomic a
        // no-op: skipped, the blocks' atomics on one line no longer queue.
        if (lo < __hip_atomic_load(&stats->lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin((long long *)&stats->lo, (long long)lo);
        if (hi > __hip_atomic_load(&stats->hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax((long long *)&stats->hi, (long long)hi);
        if (bad) atomicOr(&stats->bad, 1);
    }
}

template <class T>
__global__ __launch_bounds__(256) void to_f64_kernel(const T *__restrict__ in, int64_t n, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

// Position of row r over `dims`, first slowest (dense.rs:118-136).
__device__ __forceinline__ uint64_t position(const Dims &d, int64_t r) {
    uint64_t p = 0;
    for (int i = 0; i < d.n; i++) p = p * (uint64_t)d.extent[i] + (uint64_t)d.key[i][r];
    return p;
}

// a's rows placed into [x][z] (dense.rs:341-361): held (0 absent, 1 NULL,
// 2 valued) claimed with an exchange, so a second row at a position is a
// duplicate; reached and valued per x; finiteness of the values.
__global__ __launch_bounds__(256) void place_a_kernel(int64_t n, Dims xs, Dims zs, uint64_t nz,
                                                      const double *__restrict__ f, const uint8_t *__restrict__ fvalid,
                                                      uint32_t *held, double *a, uint8_t *reached, uint8_t *valued,
                                                      PlaceFlags *flags) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint64_t x = position(xs, r), z = position(zs, r);
    const uint64_t p = x * nz + z;
    const bool ok = fvalid[r] != 0;
    if (atomicExch(&held[p], ok ? 2u : 1u) != 0u) atomicOr(&flags->duplicate, 1);
    reached[x] = 1;
    if (ok) {
        const double v = f[r];
        a[p] = v;
        if (!isfinite(v)) atomicOr(&flags->not_finite, 1);
        if (valued) valued[x] = 1;
    }
}

// b's rows placed into [w][z][y] (dense.rs:372-386), holes left 0.  Only
// a's factors need be finite for the reference's order (dense.rs:30-35);
// b's finiteness is flagged for the MFMA contractions, which multiply every
// cell.
__global__ __launch_bounds__(256) void place_b_kernel(int64_t n, Dims wzs, Dims ys, uint64_t ny,
                                                      const double *__restrict__ f, const uint8_t *__restrict__ fvalid,
                                                      uint32_t *present, double *b, PlaceFlags *flags) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint64_t p = position(wzs, r) * ny + position(ys, r);
    if (atomicExch(&present[p], 1u) != 0u) atomicOr(&flags->duplicate, 1);
    if (fvalid[r]) {
        const double v = f[r];
        b[p] = v;
        if (!isfinite(v)) atomicOr(&flags->b_not_finite, 1);
    } else {
        atomicOr(&flags->nulls, 1);
    }
}

// Each domain row's cell of the output, or -1 (dense.rs:595-615).
__global__ __launch_bounds__(256) void locate_kernel(int64_t n, CellKeys ck, int64_t *cell) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int64_t c = 0;
    bool none = false;
    for (int i = 0; i < ck.n; i++) {
        const int64_t k = ck.key[i][r];
        const bool in = ck.valid[i][r] && k >= 0 && k < ck.extent[i];
        none |= !in;
        c += in ? k * ck.stride[i] : 0;
    }
    cell[r] = none ? -1 : c;
}

// Marks cells; a second domain row on a cell sets `repeated`.
__global__ __launch_bounds__(256) void unique_cells_kernel(int64_t n, const int64_t *cell, uint32_t *seen, int *repeated) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || cell[r] < 0) return;
    if (atomicExch(&seen[cell[r]], 1u) != 0u) atomicOr(repeated, 1);
}

// Whether every condition holds at cell (x, y) (dense.rs:479-486): keys
// are positions' digits; l <op> r + offset compared as l - r <op> offset,
// exact since keys are below 2^32.
__device__ __forceinline__ bool allowed_at(const Conds &c, uint64_t x, uint64_t y) {
    for (int i = 0; i < c.n; i++) {
        const int64_t l = (int64_t)(((c.left_in_x[i] ? x : y) / c.left_stride[i]) % c.left_extent[i]);
        const int64_t r = (int64_t)(((c.right_in_x[i] ? x : y) / c.right_stride[i]) % c.right_extent[i]);
        const int64_t d = l - r, o = c.offset[i];
        bool h;
        switch (c.op[i]) {
        case 0: h = d < o; break;
        case 1: h = d <= o; break;
        case 2: h = d > o; break;
        default: h = d >= o; break;
        }
        if (!h) return false;
    }
    return true;
}

__global__ __launch_bounds__(256) void keep_kernel(int64_t n, uint64_t ny, const uint32_t *rows, Conds c, uint32_t *keep) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    keep[o] = allowed_at(c, rows[o / ny], o % ny) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void rows_kernel(int64_t nx, const uint8_t *reached, const uint32_t *excl, uint32_t *rows) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < nx && reached[x]) rows[excl[x]] = (uint32_t)x;
}

template <class T> __device__ __forceinline__ void put(void *base, int64_t i, int64_t v) {
    static_cast<T *>(base)[i] = (T)v;
}

// The output's rows (dense.rs:516-574): for each kept cell (x, y) of the
// reached rows, its keys, value and validity.  With `keep` (masks), cell o
// of the unmasked order goes to slot excl[o].
__global__ __launch_bounds__(256) void emit_kernel(int64_t n, uint64_t ny, const uint32_t *rows,
                                                   const uint32_t *keep, const uint32_t *excl, OutKeys ok,
                                                   const double *out, const uint8_t *valued, double *values,
                                                   uint8_t *valid) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    if (keep && !keep[o]) return;
    const int64_t slot = keep ? (int64_t)excl[o] : o;
    const uint64_t x = rows[o / ny], y = o % ny;
    for (int i = 0; i < ok.n; i++) {
        const int64_t k = (int64_t)(((ok.in_x[i] ? x : y) / ok.stride[i]) % ok.extent[i]);
        switch (ok.width[i]) {
        case 1: put<int8_t>(ok.dst[i], slot, k); break;
        case 2: put<int16_t>(ok.dst[i], slot, k); break;
        case 4: put<int32_t>(ok.dst[i], slot, k); break;
        default: put<int64_t>(ok.dst[i], slot, k); break;
        }
    }
    const bool v = valued ? valued[x] != 0 : true;
    values[slot] = v ? out[x * ny + y] : 0.0;
    valid[slot] = v ? 1 : 0;
}

// A domain batch's values (dense.rs:488-515, 620-650): the cell's value
// where its group exists with one, else the default (or NULL).
__global__ __launch_bounds__(256) void domain_values_kernel(int64_t n, const int64_t *cell, uint64_t ny,
                                                            const double *out, const uint8_t *valued, Conds c,
                                                            int has_missing, double missing, double *values,
                                                            uint8_t *valid) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t cl = cell[r];
    bool got = cl >= 0;
    if (got && c.n) got = allowed_at(c, (uint64_t)cl / ny, (uint64_t)cl % ny);
    if (got && valued) got = valued[(uint64_t)cl / ny] != 0;
    if (got) {
        values[r] = out[cl];
        valid[r] = 1;
    } else {
        values[r] = has_missing ? missing : 0.0;
        valid[r] = has_missing ? 1 : 0;
    }
}

// Packs one byte per row into an Arrow validity bitmap, counting NULLs.
__global__ __launch_bounds__(256) void pack_bits_kernel(int64_t n, const uint8_t *bytes, uint8_t *bits,
                                                        unsigned long long *nulls) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = (n + 7) / 8;
    unsigned cnt = 0;
    if (i < nb) {
        uint8_t v = 0;
        for (int j = 0; j < 8; j++) {
            const int64_t r = i * 8 + j;
            if (r < n) {
                if (bytes[r]) v |= (uint8_t)(1u << j);
                else cnt++;
            }
        }
        bits[i] = v;
    }
    for (int o = 32; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(nulls, (unsigned long long)cnt);
}

__global__ __launch_bounds__(256) void u8_to_u32_kernel(int64_t n, const uint8_t *in, uint32_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// An Arrow validity bitmap (bit `offset` onward) as one byte per row.
__global__ __launch_bounds__(256) void unpack_bits_kernel(int64_t n, const uint8_t *__restrict__ bits, int64_t offset,
                                                          uint8_t *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t b = offset + r;
    out[r] = (bits[b >> 3] >> (b & 7)) & 1;
}

// A domain's key column as i64 with validity: NULLs and unsigned keys of
// 2^63 or more locate no cell (dense.rs:595-615).
template <class T>
__global__ __launch_bounds__(256) void widen_keys_kernel(int64_t n, const T *__restrict__ in, uint8_t *valid,
                                                         int64_t *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const T v = in[r];
    if constexpr (sizeof(T) == 8 && !std::is_signed<T>::value) {
        if ((uint64_t)v >> 63) { valid[r] = 0; out[r] = 0; return; }
    }
    out[r] = (int64_t)v;
}

}  // namespace

ef_status cast_keys(Type t, const void *in, int64_t n, int64_t *out, KeyStats *stats, hipStream_t s) {
    if (n == 0) return EF_OK;
    // Up to 4096 blocks (eight waves per SIMD for large columns); a block's
    // bound atomics run only when they move a bound.
    const unsigned g = blocks_for(n, 256) > 4096 ? 4096 : blocks_for(n, 256);
#define EF_CAST(T) hipLaunchKernelGGL(cast_keys_kernel<T>, dim3(g), dim3(256), 0, s, (const T *)in, n, out, stats)
    switch (t) {
    case Type::Int8: EF_CAST(int8_t); break;
    case Type::UInt8: EF_CAST(uint8_t); break;
    case Type::Int16: EF_CAST(int16_t); break;
    case Type::UInt16: EF_CAST(uint16_t); break;
    case Type::Int32: EF_CAST(int32_t); break;
    case Type::UInt32: EF_CAST(uint32_t); break;
    case Type::Int64: EF_CAST(int64_t); break;
    case Type::UInt64: EF_CAST(uint64_t); break;
    default: set_error("keys must be integers"); return EF_INVALID;
    }
#undef EF_CAST
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status to_f64(Type t, const void *in, int64_t n, double *out, hipStream_t s) {
    if (n == 0) return EF_OK;
    switch (t) {
    case Type::Float32:
        hipLaunchKernelGGL(to_f64_kernel<float>, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const float *)in, n, out);
        break;
    case Type::Float64:
        EF_HIP(hipMemcpyAsync(out, in, n * 8, hipMemcpyDeviceToDevice, s));
        break;
    default: set_error("factors must be floats"); return EF_INVALID;
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status place_a(int64_t n, const Dims &xs, const Dims &zs, uint64_t nz, const double *f, const uint8_t *fvalid,
                  uint32_t *held, double *a, uint8_t *reached, uint8_t *valued, PlaceFlags *flags, hipStream_t s) {
    if (n == 0) return EF_OK;
    hipLaunchKernelGGL(place_a_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, xs, zs, nz, f, fvalid, held, a,
                       reached, valued, flags);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status place_b(int64_t n, const Dims &wzs, const Dims &ys, uint64_t ny, const double *f, const uint8_t *fvalid,
                  uint32_t *present, double *b, PlaceFlags *flags, hipStream_t s) {
    if (n == 0) return EF_OK;
    hipLaunchKernelGGL(place_b_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, wzs, ys, ny, f, fvalid, present,
                       b, flags);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status locate(int64_t n, const CellKeys &ck, int64_t *cell, hipStream_t s) {
    if (n == 0) return EF_OK;
    hipLaunchKernelGGL(locate_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, ck, cell);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status unique_cells(int64_t n, const int64_t *cell, uint32_t *seen, int *repeated, hipStream_t s) {
    if (n == 0) return EF_OK;
    hipLaunchKernelGGL(unique_cells_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, cell, seen, repeated);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status reached_rows(int64_t nx, const uint8_t *reached, uint32_t *tmp32, uint32_t *excl, uint32_t *rows,
                       void *scan_tmp, hipStream_t s) {
    if (nx == 0) return EF_OK;
    hipLaunchKernelGGL(u8_to_u32_kernel, dim3(blocks_for(nx, 256)), dim3(256), 0, s, nx, reached, tmp32);
    EF_TRY(scan_u32(tmp32, excl, (size_t)nx, scan_tmp, s));
    hipLaunchKernelGGL(rows_kernel, dim3(blocks_for(nx, 256)), dim3(256), 0, s, nx, reached, excl, rows);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status keep_cells(int64_t n, uint64_t ny, const uint32_t *rows, const Conds &c, uint32_t *keep, hipStream_t s) {
    if (n == 0) return EF_OK;
    hipLaunchKernelGGL(keep_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, ny, rows, c, keep);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status emit(int64_t n, uint64_t ny, const uint32_t *rows, const uint32_t *keep, const uint32_t *excl,
               const OutKeys &ok, const double *out, const uint8_t *valued, double *values, uint8_t *valid,
               hipStream_t s) {
    if (n == 0) return EF_OK;
    hipLaunchKernelGGL(emit_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, ny, rows, keep, excl, ok, out,
                       valued, values, valid);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status domain_values(int64_t n, const int64_t *cell, uint64_t ny, const double *out, const uint8_t *valued,
                        const Conds &c, int has_missing, double missing, double *values, uint8_t *valid,
                        hipStream_t s) {
    if (n == 0) return EF_OK;
    hipLaunchKernelGGL(domain_values_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, cell, ny, out, valued, c,
                       has_missing, missing, values, valid);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status pack_bits(int64_t n, const uint8_t *bytes, uint8_t *bits, unsigned long long *nulls, hipStream_t s) {
    if (n == 0) return EF_OK;
    hipLaunchKernelGGL(pack_bits_kernel, dim3(blocks_for((n + 7) / 8, 256)), dim3(256), 0, s, n, bytes, bits, nulls);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status unpack_bits(int64_t n, const uint8_t *bits, int64_t offset, uint8_t *out, hipStream_t s) {
    if (n == 0) return EF_OK;
    hipLaunchKernelGGL(unpack_bits_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, bits, offset, out);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status widen_keys(Type t, int64_t n, const void *in, uint8_t *valid, int64_t *out, hipStream_t s) {
    if (n == 0) return EF_OK;
    const unsigned g = blocks_for(n, 256);
#define EF_W(T) hipLaunchKernelGGL(widen_keys_kernel<T>, dim3(g), dim3(256), 0, s, n, (const T *)in, valid, out)
    switch (t) {
    case Type::Int8: EF_W(int8_t); break;
    case Type::UInt8: EF_W(uint8_t); break;
    case Type::Int16: EF_W(int16_t); break;
    case Type::UInt16: EF_W(uint16_t); break;
    case Type::Int32: EF_W(int32_t); break;
    case Type::UInt32: EF_W(uint32_t); break;
    case Type::Int64: EF_W(int64_t); break;
    case Type::UInt64: EF_W(uint64_t); break;
    default: set_error("a domain's keys must be integers"); return EF_INVALID;
    }
#undef EF_W
    EF_HIP(hipGetLastError());
    return EF_OK;
}

}  // namespace ef
