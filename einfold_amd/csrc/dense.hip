// The dense positional fold's contraction on gfx950
// (crates/einfold-kernels/src/dense.rs:706-770), and batched matrix
// products on MFMA tiles.
//
// The contraction reproduces the reference's summation exactly: each output
// cell starts at +0.0 and adds a[x][z] * b[w][z][y] for z in order over the
// cells `held` marks valued, rounding the product and the sum separately
// (rustc never contracts to FMA).  Blocks of output cells run in parallel,
// but each cell's chain is sequential, so the bits are the reference's.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

#include "common.h"
#include "mfma.h"

namespace ef {

namespace {

constexpr uint32_t VALUED = 2;
constexpr int ZC = 16;   // z positions staged per step

// Output tile (16*RM) x (16*RN) per workgroup of 256 threads, for rows of
// one batch.  PRED: skip non-valued cells with a select (exact whatever b
// holds); without it, a non-valued cell's factor is 0, exact when every b
// is finite (0 * b is then +-0, and a sum from +0.0 is never -0.0).
template <int RM, int RN, bool PRED>
__global__ __launch_bounds__(256) void dense_accumulate_kernel(const double *__restrict__ a,
                                                               const uint32_t *__restrict__ held,
                                                               const double *__restrict__ b,
                                                               int64_t nx, int64_t nz, int64_t ny,
                                                               int64_t per_batch, int64_t x_tiles,
                                                               double *__restrict__ out) {
#pragma clang fp contract(off)
    constexpr int TM = 16 * RM, TN = 16 * RN;
    __shared__ double as[TM][ZC + 1];
    __shared__ double bs[ZC][TN];
    __shared__ uint8_t fs[TM][ZC + 1];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int64_t w = blockIdx.z;
    const int64_t xt = blockIdx.x;
    const int64_t x0 = w * per_batch + xt * TM;          // first row (global)
    const int64_t xend = w * per_batch + per_batch;      // end of this batch's rows
    const int64_t y0 = (int64_t)blockIdx.y * TN;
    (void)x_tiles;
    double acc[RM][RN];
#pragma unroll
    for (int r = 0; r < RM; r++)
#pragma unroll
        for (int c = 0; c < RN; c++) acc[r][c] = 0.0;
    const double *bw = b + w * nz * ny;
    for (int64_t z0 = 0; z0 < nz; z0 += ZC) {
        static_assert(TM * ZC % 256 == 0 && ZC * TN % 256 == 0, "whole staging rounds");
        constexpr int NA = TM * ZC / 256, NB = ZC * TN / 256;
        // Every load of the step issued before any LDS store, each from a
        // clamped cell (x0 and z0 are inside), then 0 outside: a load behind
        // `in &&` becomes a branch and a wait per element.
        double ra[NA], rb[NB];
        bool rv[NA];
#pragma unroll
        for (int k = 0; k < NA; k++) {
            const int i = tid + 256 * k, r = i / ZC, zz = i % ZC;
            const int64_t x = x0 + r, z = z0 + zz;
            const bool in = x < xend && x < nx && z < nz;
            const int64_t xc = in ? x : x0, zc = in ? z : z0;
            ra[k] = a[xc * nz + zc];
            const uint32_t hv = held[xc * nz + zc];
            rv[k] = in && hv == VALUED;
        }
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const int i = tid + 256 * k, zz = i / TN, cc = i % TN;
            const int64_t z = z0 + zz, y = y0 + cc;
            const bool inb = z < nz && y < ny;
            const double bv = bw[(inb ? z : z0) * ny + (inb ? y : y0)];
            rb[k] = inb ? bv : 0.0;
        }
#pragma unroll
        for (int k = 0; k < NA; k++) {
            const int i = tid + 256 * k;
            as[i / ZC][i % ZC] = rv[k] ? ra[k] : 0.0;
            fs[i / ZC][i % ZC] = rv[k] ? 1 : 0;
        }
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const int i = tid + 256 * k;
            bs[i / TN][i % TN] = rb[k];
        }
        __syncthreads();
        const int zn = nz - z0 < ZC ? (int)(nz - z0) : ZC;
        for (int zz = 0; zz < zn; zz++) {
            double bv[RN];
#pragma unroll
            for (int c = 0; c < RN; c++) bv[c] = bs[zz][tx + 16 * c];
#pragma unroll
            for (int r = 0; r < RM; r++) {
                const double av = as[ty + 16 * r][zz];
                if (PRED) {
                    const bool f = fs[ty + 16 * r][zz] != 0;
#pragma unroll
                    for (int c = 0; c < RN; c++) {
                        const double t = acc[r][c] + av * bv[c];
                        acc[r][c] = f ? t : acc[r][c];
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < RN; c++) acc[r][c] = acc[r][c] + av * bv[c];
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < RM; r++) {
        const int64_t x = x0 + ty + 16 * r;
        if (x >= xend || x >= nx) continue;
#pragma unroll
        for (int c = 0; c < RN; c++) {
            const int64_t y = y0 + tx + 16 * c;
            if (y < ny) out[x * ny + y] = acc[r][c];
        }
    }
}

// Outputs too small to fill the chip with the tiles above (256^3 makes 64
// workgroups of 32 x 32): one cell per thread, 16 x 16 cells per
// workgroup, z in chunks of 256 staged through LDS (69 KB), the next chunk
// loaded into registers while this one is summed.  A chunk's chain of 256
// sums lasts about one memory latency, so the next chunk's loads are hidden
// behind it, and 256^3 takes one exposed load in all.  Each cell's chain is
// the same z-ordered sequence of roundings, so the same bits.  PRED: the
// caller does not vouch that b is finite, so each chunk checks its b slab
// and skips non-valued cells by their flags where the slab is not.
template <bool PRED>
__global__ __launch_bounds__(256) void dense_accumulate_thin(const double *__restrict__ a,
                                                             const uint32_t *__restrict__ held,
                                                             const double *__restrict__ b, int64_t nx, int64_t nz,
                                                             int64_t ny, int64_t per_batch, double *__restrict__ out) {
#pragma clang fp contract(off)
    constexpr int T = 16, ZT = 256, PER = T * ZT / 256;   // 16 values of a and of b per thread per chunk
    __shared__ double as[T][ZT + 1];
    __shared__ uint8_t fs[T][ZT + 1];
    __shared__ double bs[ZT][T];
    __shared__ int wave_finite[4];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4, lane = tid & 63;
    const int64_t w = blockIdx.z;
    const int64_t x0 = w * per_batch + (int64_t)blockIdx.x * T;
    const int64_t xend = w * per_batch + per_batch < nx ? w * per_batch + per_batch : nx;
    const int64_t y0 = (int64_t)blockIdx.y * T;
    const double *bw = b + w * nz * ny;
    double ra[PER], rb[PER];
    uint8_t rf[PER];
    auto fetch = [&](int64_t z0) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + 256 * k;
            const int r = i / ZT, zz = i % ZT;
            const int64_t x = x0 + r, z = z0 + zz;
            const bool in = x < xend && z < nz;
            // Unconditional loads from a clamped cell, then 0 outside.
            const int64_t xc = in ? x : x0, zcl = in ? z : 0;
            const double av = a[xc * nz + zcl];
            // The flag read unconditionally too: behind `in &&` it becomes a
            // branch with a wait, and the 16 loads run one after another.
            const uint32_t hv = held[xc * nz + zcl];
            const bool v = in && hv == VALUED;
            ra[k] = v ? av : 0.0;
            rf[k] = v ? 1 : 0;
            const int zb = i / T, cc = i % T;
            const int64_t zb_g = z0 + zb, y = y0 + cc;
            const bool inb = zb_g < nz && y < ny;
            const double bv = bw[(inb ? zb_g : 0) * ny + (inb ? y : 0)];
            rb[k] = inb ? bv : 0.0;
        }
    };
    double acc = 0.0;
    // One chunk's sums.  FLAGS: skip the cells `held` does not mark valued
    // with a select on the product; without, their zero factors give +-0
    // products, which leave a sum from +0.0 as it is (such a sum is never
    // -0.0) when b is finite: the same bits either way.
    auto sum_chunk = [&](auto flags, int zn) {
        constexpr bool FLAGS = decltype(flags)::value;
        // 16 z at a time: every LDS read of the run issued before the chain
        // of sums that uses them, so the chain waits on one read latency per
        // run instead of one per z.
        int zz0 = 0;
        for (; zz0 + 16 <= zn; zz0 += 16) {
            double av[16], bv[16];
            bool fv[16];
#pragma unroll
            for (int j = 0; j < 16; j++) {
                av[j] = as[ty][zz0 + j];
                bv[j] = bs[zz0 + j][tx];
                fv[j] = !FLAGS || fs[ty][zz0 + j];
            }
            // A skipped cell adds -0.0, which leaves every sum as it is
            // (x + -0.0 == x, signed zeros included): the select is taken on
            // the product, off the chain of dependent sums.
#pragma unroll
            for (int j = 0; j < 16; j++) acc = acc + (fv[j] ? av[j] * bv[j] : -0.0);
        }
        for (int zz = zz0; zz < zn; zz++) acc = acc + ((!FLAGS || fs[ty][zz]) ? as[ty][zz] * bs[zz][tx] : -0.0);
    };
    if (nz > 0) fetch(0);
    for (int64_t z0 = 0; z0 < nz; z0 += ZT) {
        int finite = 1;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int i = tid + 256 * k;
            as[i / ZT][i % ZT] = ra[k];
            fs[i / ZT][i % ZT] = rf[k];
            bs[i / T][i % T] = rb[k];
            finite &= isfinite(rb[k]) ? 1 : 0;
        }
        // The flags are needed only where the workgroup's b slab of this
        // chunk holds an infinite or NaN value: a vote per wave, read by all
        // after the barrier that also publishes the tiles (each wave's slot
        // is written again only after the chunk's closing barrier).
        const int wave_all = __all(finite);   // the vote of every lane, outside the branch
        if (PRED && lane == 0) wave_finite[tid >> 6] = wave_all;
        __syncthreads();
        const bool flags = PRED && !(wave_finite[0] && wave_finite[1] && wave_finite[2] && wave_finite[3]);
        if (z0 + ZT < nz) fetch(z0 + ZT);
        const int zn = nz - z0 < ZT ? (int)(nz - z0) : ZT;
        if (flags) sum_chunk(std::true_type{}, zn);
        else sum_chunk(std::false_type{}, zn);
        __syncthreads();
    }
    const int64_t x = x0 + ty, y = y0 + tx;
    if (x < xend && y < ny) out[x * ny + y] = acc;
}

// Outputs of fewer than 4096 cells (dense.rs:752-768): z in blocks of
// 4096, each block's partial summed from +0.0, the partials added to +0.0
// in block order.  One thread per cell.
__global__ __launch_bounds__(256) void dense_accumulate_small(const double *__restrict__ a,
                                                              const uint32_t *__restrict__ held,
                                                              const double *__restrict__ b,
                                                              int64_t nx, int64_t nz, int64_t ny,
                                                              int64_t per_batch,
                                                              double *__restrict__ out) {
#pragma clang fp contract(off)
    constexpr int64_t ZB = 4096;
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= nx * ny) return;
    const int64_t x = cell / ny, y = cell % ny;
    const double *bw = b + (x / per_batch) * nz * ny;
    double total = 0.0;
    for (int64_t z0 = 0; z0 < nz; z0 += ZB) {
        const int64_t z1 = z0 + ZB < nz ? z0 + ZB : nz;
        double part = 0.0;
        for (int64_t z = z0; z < z1; z++) {
            if (held[x * nz + z] == VALUED) part = part + a[x * nz + z] * bw[z * ny + y];
        }
        total = total + part;
    }
    out[cell] = total;
}

// Batched C[w] = A[w] (m x k) B[w] (k x n), row-major, on MFMA tiles: a
// 64 x 64 output tile per workgroup, 32 x 32 per wave, K in steps of 32.
template <int P>
__global__ __launch_bounds__(256) void product_kernel(const typename Input<P>::T *__restrict__ A,
                                                      const typename Input<P>::T *__restrict__ B,
                                                      typename Input<P>::T *__restrict__ C,
                                                      int64_t m, int64_t n, int64_t k) {
    typedef typename Input<P>::T In;
    typedef typename Acc<P>::type AccV;
    typedef typename Elem<P>::T E;
    constexpr int T = 64, KS = 32, PITCH = KS + 8, PLANE = T * PITCH;
    __shared__ __attribute__((aligned(16))) E as[Elem<P>::planes * PLANE];
    __shared__ __attribute__((aligned(16))) E bts[Elem<P>::planes * PLANE];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, c = lane & 15;
    const int64_t bz = blockIdx.z;
    const int64_t m0 = (int64_t)blockIdx.y * T, n0 = (int64_t)blockIdx.x * T;
    const In *Aw = A + bz * m * k;
    const In *Bw = B + bz * k * n;
    In *Cw = C + bz * m * n;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    AccV acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = AccV{0, 0, 0, 0};
    for (int64_t k0 = 0; k0 < k; k0 += KS) {
        // A tile [64 rows][32 k]; B^T tile [64 cols][32 k].
        for (int i = tid; i < T * KS; i += 256) {
            const int r = i / KS, kk = i % KS;
            const int64_t row = m0 + r, kx = k0 + kk;
            tile_store<P>(as, PLANE, r * PITCH + kk, (row < m && kx < k) ? Aw[row * k + kx] : In(0));
        }
        for (int i = tid; i < T * KS; i += 256) {
            const int kk = i / T, cc = i % T;
            const int64_t col = n0 + cc, kx = k0 + kk;
            tile_store<P>(bts, PLANE, cc * PITCH + kk, (col < n && kx < k) ? Bw[kx * n + col] : In(0));
        }
        __syncthreads();
        Frag<P> af[2], bf[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            tile_load8<P>(af[i], as, PLANE, (wm + i * 16 + c) * PITCH + g * 8);
            tile_load8<P>(bf[i], bts, PLANE, (wn + i * 16 + c) * PITCH + g * 8);
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) mma<P>(acc[i][j], af[i], bf[j]);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wm + i * 16 + crow<P>(g, r), col = n0 + wn + j * 16 + c;
                if (row < m && col < n) Cw[row * n + col] = (In)acc[i][j][r];
            }
}

template <int P>
ef_status launch_product(const typename Input<P>::T *a, const typename Input<P>::T *b,
                         typename Input<P>::T *out, size_t batches, size_t m, size_t n, size_t k,
                         hipStream_t s) {
    if (batches == 0 || m == 0 || n == 0) return EF_OK;
    if (batches > 65535) {
        set_error("at most 65535 products per call");
        return EF_INVALID;
    }
    dim3 grid((unsigned)((n + 63) / 64), (unsigned)((m + 63) / 64), (unsigned)batches);
    hipLaunchKernelGGL(product_kernel<P>, grid, dim3(256), 0, s, a, b, out, (int64_t)m, (int64_t)n, (int64_t)k);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

}  // namespace

ef_status launch_dense_accumulate_f64(const double *a, const uint32_t *held, const double *b,
                                      size_t nx, size_t nz, size_t ny, size_t per_batch,
                                      int predicated, double *out, hipStream_t s) {
    if (nx == 0 || ny == 0) return EF_OK;
    if (per_batch == 0 || nx % per_batch != 0) {
        set_error("rows must be whole batches of per_batch rows");
        return EF_INVALID;
    }
    const int64_t nw = (int64_t)(nx / per_batch);
    if (nx * ny < 4096) {
        hipLaunchKernelGGL(dense_accumulate_small, dim3(blocks_for(nx * ny, 256)), dim3(256), 0, s,
                           a, held, b, (int64_t)nx, (int64_t)nz, (int64_t)ny, (int64_t)per_batch, out);
    } else {
        // 64 x 64 tiles when they fill the chip, else 32 x 32, else (fewer
        // than 512 of those) one cell per thread.
        const bool big = ((per_batch + 63) / 64) * ((ny + 63) / 64) * (size_t)nw >= 1024;
        const bool thin = ((per_batch + 31) / 32) * ((ny + 31) / 32) * (size_t)nw < 512;
        if (thin) {
            if (nw > 65535 || (ny + 15) / 16 > 65535 || (per_batch + 15) / 16 > 0x7fffffffULL) {
                set_error("too many batches or columns for one launch");
                return EF_INVALID;
            }
            dim3 tgrid((unsigned)((per_batch + 15) / 16), (unsigned)((ny + 15) / 16), (unsigned)nw);
            if (predicated)
                hipLaunchKernelGGL(dense_accumulate_thin<true>, tgrid, dim3(256), 0, s, a, held, b, (int64_t)nx,
                                   (int64_t)nz, (int64_t)ny, (int64_t)per_batch, out);
            else
                hipLaunchKernelGGL(dense_accumulate_thin<false>, tgrid, dim3(256), 0, s, a, held, b, (int64_t)nx,
                                   (int64_t)nz, (int64_t)ny, (int64_t)per_batch, out);
            EF_HIP(hipGetLastError());
            return EF_OK;
        }
        const int tm = big ? 64 : 32, tn = big ? 64 : 32;
        const int64_t x_tiles = ((int64_t)per_batch + tm - 1) / tm;
        if (nw > 65535 || (ny + tn - 1) / tn > 65535) {
            set_error("too many batches or columns for one launch");
            return EF_INVALID;
        }
        dim3 grid((unsigned)x_tiles, (unsigned)((ny + tn - 1) / tn), (unsigned)nw);
#define EF_ACC(RM, RN, PR)                                                                   \
    hipLaunchKernelGGL((dense_accumulate_kernel<RM, RN, PR>), grid, dim3(256), 0, s, a, held, b, \
                       (int64_t)nx, (int64_t)nz, (int64_t)ny, (int64_t)per_batch, x_tiles, out)
        if (big) {
            if (predicated) EF_ACC(4, 4, true); else EF_ACC(4, 4, false);
        } else {
            if (predicated) EF_ACC(2, 2, true); else EF_ACC(2, 2, false);
        }
#undef EF_ACC
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status launch_dense_product_f32(const float *a, const float *b, float *out, size_t batches,
                                   size_t m, size_t n, size_t k, uint32_t precision, hipStream_t s) {
    switch (precision) {
    case EF_PRECISION_FP32: return launch_product<PREC_F32>(a, b, out, batches, m, n, k, s);
    case EF_PRECISION_BF16X3: return launch_product<PREC_BF16X3>(a, b, out, batches, m, n, k, s);
    case EF_PRECISION_BF16: return launch_product<PREC_BF16>(a, b, out, batches, m, n, k, s);
    default:
        set_error("FP32 products take precision FP32, BF16X3 or BF16");
        return EF_INVALID;
    }
}

ef_status launch_dense_product_f64(const double *a, const double *b, double *out, size_t batches,
                                   size_t m, size_t n, size_t k, hipStream_t s) {
    return launch_product<PREC_F64>(a, b, out, batches, m, n, k, s);
}

}  // namespace ef
