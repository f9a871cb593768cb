// Wave-level 16x16 MFMA fragments for gfx950, in one logical layout for
// every precision.
//
// A 16x16 output block C[m][n] += sum_k A[m][k] B[k][n] is computed over
// K in chunks of 32.  In each chunk, lane l (group g = l / 16) holds 8
// logical k-values of its row m = l % 16 of A and of its column n = l % 16
// of B: logical k = g*8 + j, j = 0..7.  The physical k a logical k stands
// for is the caller's choice, as long as A and B agree:
//   * contiguous:  physical k = g*8 + j            (operands read from memory)
//   * cperm:       physical k = crow(g, j) (j < 4), 16 + crow(g, j-4) (j >= 4)
//                  (an operand taken from two stacked C blocks in registers)
// The C block's layout: lane l holds C[crow<P>(g, i)][l % 16], i = 0..3,
// where crow is g*4 + i for the FP32 and BF16 instructions and 4*i + g for
// v_mfma_f64_16x16x4_f64 (its rows interleave across the lane groups).
//
// Precisions: FP32 issues eight v_mfma_f32_16x16x4_f32 per chunk (step t
// takes logical k = g*8 + t from group g); FP64 likewise with
// v_mfma_f64_16x16x4_f64; BF16 one v_mfma_f32_16x16x32_bf16; BF16X3 three
// of them on split operands x = hi + lo (hi = bf16(x), lo = bf16(x - hi)),
// hi*hi + hi*lo + lo*hi, recovering about 16 more bits than one BF16
// product.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ef {

enum : int { PREC_F64 = 0, PREC_F32 = 1, PREC_BF16X3 = 2, PREC_BF16 = 3 };

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <int P> struct Frag;
template <> struct Frag<PREC_F32> { float v[8]; };
template <> struct Frag<PREC_F64> { double v[8]; };
template <> struct Frag<PREC_BF16> { bf16x8 hi; };
template <> struct Frag<PREC_BF16X3> { bf16x8 hi, lo; };

// Row of a 16x16 C block held in register i of lane group g.
template <int P> __device__ __forceinline__ constexpr int crow(int g, int i) {
    return P == PREC_F64 ? 4 * i + g : 4 * g + i;
}

template <int P> struct Acc { typedef f32x4 type; typedef float scalar; };
template <> struct Acc<PREC_F64> { typedef f64x4 type; typedef double scalar; };

// Element type of a precision's tiles in LDS, and how many planes (hi, lo).
template <int P> struct Elem { typedef float T; static constexpr int planes = 1; };
template <> struct Elem<PREC_F64> { typedef double T; static constexpr int planes = 1; };
template <> struct Elem<PREC_BF16> { typedef __bf16 T; static constexpr int planes = 1; };
template <> struct Elem<PREC_BF16X3> { typedef __bf16 T; static constexpr int planes = 2; };

template <int P> struct Input { typedef float T; };
template <> struct Input<PREC_F64> { typedef double T; };

__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

template <int P> __device__ __forceinline__ void mma(typename Acc<P>::type &c, const Frag<P> &a, const Frag<P> &b);

template <> __device__ __forceinline__ void mma<PREC_F32>(f32x4 &c, const Frag<PREC_F32> &a, const Frag<PREC_F32> &b) {
#pragma unroll
    for (int t = 0; t < 8; t++) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.v[t], b.v[t], c, 0, 0, 0);
}
template <> __device__ __forceinline__ void mma<PREC_F64>(f64x4 &c, const Frag<PREC_F64> &a, const Frag<PREC_F64> &b) {
#pragma unroll
    for (int t = 0; t < 8; t++) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a.v[t], b.v[t], c, 0, 0, 0);
}
template <> __device__ __forceinline__ void mma<PREC_BF16>(f32x4 &c, const Frag<PREC_BF16> &a, const Frag<PREC_BF16> &b) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
}
template <> __device__ __forceinline__ void mma<PREC_BF16X3>(f32x4 &c, const Frag<PREC_BF16X3> &a, const Frag<PREC_BF16X3> &b) {
    // Small products first, so the large one is added to their sum.
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
}

// One product as separately issued sub-instructions, so that a caller can
// interleave independent accumulators between them: BF16x3 has three
// (lo*hi, hi*lo, hi*hi), BF16 one, FP32 and FP64 eight k-steps.
template <int P> struct Parts { static constexpr int n = 8; };
template <> struct Parts<PREC_BF16X3> { static constexpr int n = 3; };
template <> struct Parts<PREC_BF16> { static constexpr int n = 1; };

template <int P>
__device__ __forceinline__ void mma_part(typename Acc<P>::type &c, const Frag<P> &a, const Frag<P> &b, int s) {
    if constexpr (P == PREC_F32) {
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.v[s], b.v[s], c, 0, 0, 0);
    } else if constexpr (P == PREC_F64) {
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a.v[s], b.v[s], c, 0, 0, 0);
    } else if constexpr (P == PREC_BF16) {
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
    } else {
        if (s == 0) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, c, 0, 0, 0);
        else if (s == 1) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, c, 0, 0, 0);
        else c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
    }
}

// 2^x: the hardware's v_exp_f32 for float (no denormal results: a weight
// below 2^-126 becomes 0), the library's exp2 for double.
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ double fast_exp2(double x) { return exp2(x); }

// Slots j and j + 1 (j even) of a fragment from two values (split for
// BF16 precisions: two values per conversion, hi read back from the packed
// bits).
template <int P, class T> __device__ __forceinline__ void set_pair(Frag<P> &f, int j, T y0, T y1) {
    if constexpr (P == PREC_F32) {
        f.v[j] = (float)y0;
        f.v[j + 1] = (float)y1;
    } else if constexpr (P == PREC_F64) {
        f.v[j] = (double)y0;
        f.v[j + 1] = (double)y1;
    } else if constexpr (P == PREC_BF16) {
        f.hi[j] = (__bf16)(float)y0;
        f.hi[j + 1] = (__bf16)(float)y1;
    } else {
        const float x0 = (float)y0, x1 = (float)y1;
        const bf16x2 h = __builtin_convertvector((f32x2){x0, x1}, bf16x2);
        const unsigned u = __builtin_bit_cast(unsigned, h);
        const f32x2 r = {x0 - __uint_as_float(u << 16), x1 - __uint_as_float(u & 0xffff0000u)};
        const bf16x2 l = __builtin_convertvector(r, bf16x2);
        f.hi[j] = h[0];
        f.hi[j + 1] = h[1];
        f.lo[j] = l[0];
        f.lo[j + 1] = l[1];
    }
}

// A fragment from 8 values in registers.
template <int P, class T> __device__ __forceinline__ void make_frag(Frag<P> &f, const T (&x)[8]) {
#pragma unroll
    for (int j = 0; j < 8; j += 2) set_pair<P>(f, j, x[j], x[j + 1]);
}

// Stores one value into a tile of `plane` elements per plane.
template <int P, class T>
__device__ __forceinline__ void tile_store(typename Elem<P>::T *base, int plane, int idx, T x) {
    if constexpr (P == PREC_BF16X3) {
        __bf16 h, l;
        split_bf16((float)x, h, l);
        base[idx] = h;
        base[plane + idx] = l;
    } else {
        base[idx] = (typename Elem<P>::T)x;
    }
}

// Loads logical k = 0..7 from base[idx .. idx+7] (idx 8-element aligned).
template <int P>
__device__ __forceinline__ void tile_load8(Frag<P> &f, const typename Elem<P>::T *base, int plane, int idx) {
    if constexpr (P == PREC_F32) {
        const f32x4 a = *reinterpret_cast<const f32x4 *>(base + idx);
        const f32x4 b = *reinterpret_cast<const f32x4 *>(base + idx + 4);
#pragma unroll
        for (int j = 0; j < 4; j++) { f.v[j] = a[j]; f.v[j + 4] = b[j]; }
    } else if constexpr (P == PREC_F64) {
#pragma unroll
        for (int j = 0; j < 8; j++) f.v[j] = base[idx + j];
    } else if constexpr (P == PREC_BF16) {
        f.hi = *reinterpret_cast<const bf16x8 *>(base + idx);
    } else {
        f.hi = *reinterpret_cast<const bf16x8 *>(base + idx);
        f.lo = *reinterpret_cast<const bf16x8 *>(base + plane + idx);
    }
}

// Loads the cperm pattern of one 32-wide chunk starting at base[i]:
// logical k = j holds physical crow(g, j) (j < 4) and 16 + crow(g, j - 4)
// (j >= 4).  FP32/BF16: two runs of 4 (4-element aligned); FP64: strided.
template <int P>
__device__ __forceinline__ void tile_load_cperm(Frag<P> &f, const typename Elem<P>::T *base, int plane, int i, int g) {
    if constexpr (P == PREC_F64) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            f.v[j] = base[i + crow<P>(g, j)];
            f.v[j + 4] = base[i + 16 + crow<P>(g, j)];
        }
    } else if constexpr (P == PREC_F32) {
        const f32x4 a = *reinterpret_cast<const f32x4 *>(base + i + g * 4);
        const f32x4 b = *reinterpret_cast<const f32x4 *>(base + i + 16 + g * 4);
#pragma unroll
        for (int j = 0; j < 4; j++) { f.v[j] = a[j]; f.v[j + 4] = b[j]; }
    } else {
        const bf16x4 a = *reinterpret_cast<const bf16x4 *>(base + i + g * 4);
        const bf16x4 b = *reinterpret_cast<const bf16x4 *>(base + i + 16 + g * 4);
#pragma unroll
        for (int j = 0; j < 4; j++) { f.hi[j] = a[j]; f.hi[j + 4] = b[j]; }
        if constexpr (P == PREC_BF16X3) {
            const bf16x4 c = *reinterpret_cast<const bf16x4 *>(base + plane + i + g * 4);
            const bf16x4 d = *reinterpret_cast<const bf16x4 *>(base + plane + i + 16 + g * 4);
#pragma unroll
            for (int j = 0; j < 4; j++) { f.lo[j] = c[j]; f.lo[j + 4] = d[j]; }
        }
    }
}

// The values of lanes i % 32 and 32 + i % 32 (a, b), and of lanes
// 32 * (i / 32) + i % 16 and that + 16 (rows 0/1 or 2/3 of the wave), by
// gfx950's lane-swap instructions: no LDS round trip.
__device__ __forceinline__ void lanes32(float x, float &a, float &b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void lanes16(float x, float &a, float &b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void lanes32(double x, double &a, double &b) {
    const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
    const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    a = __hiloint2double((int)rh[0], (int)rl[0]);
    b = __hiloint2double((int)rh[1], (int)rl[1]);
}
__device__ __forceinline__ void lanes16(double x, double &a, double &b) {
    const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
    const auto rl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto rh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    a = __hiloint2double((int)rh[0], (int)rl[0]);
    b = __hiloint2double((int)rh[1], (int)rl[1]);
}

// Max over the four lane groups that share a column (l % 16), the same on
// every lane of the column (NaN is ignored unless all four are NaN).
template <class T> __device__ __forceinline__ T wave_max(T x) {
    T a, b;
    lanes16(x, a, b);
    x = fmax(a, b);
    lanes32(x, a, b);
    return fmax(a, b);
}

// Sum over the four lane groups that share a column, in a fixed order:
// (g0 + g1) + (g2 + g3) on every lane (addition commutes, so each lane's
// pairs give the same bits).
template <class T> __device__ __forceinline__ T group_sum(T x) {
    T a, b;
    lanes16(x, a, b);
    x = a + b;
    lanes32(x, a, b);
    return a + b;
}

}  // namespace ef
