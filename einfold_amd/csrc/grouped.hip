// Many uneven dense products in one persistent launch: einfold's batched
// dense folds (dense.rs:706-770 with Float32 factors, products C = A B per
// fold) as a grouped GEMM on MFMA.
//
// A prep launch writes every problem's operands once in the form the
// fragments read: A as [plane][m][k] and B transposed as [plane][n][k]
// (k contiguous in both), padded with zeros to whole tiles, split into BF16
// hi and lo planes for BF16x3.  The GEMM launch is persistent: workgroups
// take (problem, tile row, tile column) items from a counter, largest K
// first, and compute each output tile whole, k in order, so the bits of a
// tile never depend on which workgroup took it or on the other problems.
// Tiles stream through double-buffered LDS with the next k step held in
// registers while the current one is multiplied.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.h"
#include "mfma.h"

namespace ef {

namespace {

constexpr int KS = 64;          // k per step
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

struct GemmDesc {
    const float *a, *b;
    float *c;
    int64_t m, n, k;
    int64_t mp, np, kp;          // padded
    int64_t a_off, b_off;        // element offsets of the prepared planes in the operand buffer
};

struct PrepTask {
    int problem, which;          // which 0: A, 1: B
    int bi, bj;                  // 64 x 64 block of the destination (rows, k)
};

struct Item {
    int problem, tm, tn, pad;
};


// Eight consecutive k of one row of a prepared operand, in the element
// type of precision P: two planes (hi, lo) for BF16x3.  dst is 16-byte
// aligned (k is padded to whole 64-element steps).
template <int P>
__device__ __forceinline__ void put8(typename Elem<P>::T *dst, int64_t plane, const float (&v)[8]) {
    if constexpr (P == PREC_BF16X3) {
        bf16x8 h, l;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            __bf16 hh, ll;
            split_bf16(v[j], hh, ll);
            h[j] = hh;
            l[j] = ll;
        }
        *reinterpret_cast<bf16x8 *>(dst) = h;
        *reinterpret_cast<bf16x8 *>(dst + plane) = l;
    } else if constexpr (P == PREC_BF16) {
        bf16x8 h;
#pragma unroll
        for (int j = 0; j < 8; j++) h[j] = (__bf16)v[j];
        *reinterpret_cast<bf16x8 *>(dst) = h;
    } else {
        *reinterpret_cast<f32x4 *>(dst) = f32x4{v[0], v[1], v[2], v[3]};
        *reinterpret_cast<f32x4 *>(dst + 4) = f32x4{v[4], v[5], v[6], v[7]};
    }
}

typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));

// Writes one 64 x 64 block of a prepared operand: destination rows r (of m
// or n) and columns kk (of k), as 8-element chunks (two per thread).  A is
// read as [m][k]; B as [k][n] and transposed through LDS.  Cells outside
// the matrix are zeros.
template <int P>
__global__ __launch_bounds__(256) void grouped_prep_kernel(const GemmDesc *descs, const PrepTask *tasks,
                                                           typename Elem<P>::T *ops) {
    typedef typename Elem<P>::T E;
    __shared__ float t[64][65];
    const PrepTask task = tasks[blockIdx.x];
    const GemmDesc d = descs[task.problem];
    const int64_t r0 = (int64_t)task.bi * 64, k0 = (int64_t)task.bj * 64;
    const int tid = threadIdx.x;
    float v[2][8];
    if (task.which == 0) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int q = tid + 256 * i;
            const int64_t row = r0 + q / 8, col = k0 + (q % 8) * 8;
            if (row >= d.m) {
#pragma unroll
                for (int j = 0; j < 8; j++) v[i][j] = 0.f;
            } else if (col + 8 <= d.k) {
                const float *src = d.a + row * d.k + col;
                const f32x4u x0 = *reinterpret_cast<const f32x4u *>(src);
                const f32x4u x1 = *reinterpret_cast<const f32x4u *>(src + 4);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    v[i][j] = x0[j];
                    v[i][j + 4] = x1[j];
                }
            } else {
                // Unconditional loads from clamped cells, then 0 outside:
                // the loads issue together instead of one branch and wait
                // each.
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const float a = d.a[row * d.k + (col + j < d.k ? col + j : d.k - 1)];
                    v[i][j] = col + j < d.k ? a : 0.f;
                }
            }
        }
    } else {
        // Read B[k0 + kk][r0 + r] (coalesced along n) into t[kk][r].
        for (int i = tid; i < 64 * 64; i += 256) {
            const int kk = i / 64, r = i % 64;
            const int64_t kx = k0 + kk, col = r0 + r;
            const float b = d.b[(kx < d.k ? kx : d.k - 1) * d.n + (col < d.n ? col : d.n - 1)];
            t[kk][r] = (kx < d.k && col < d.n) ? b : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int q = tid + 256 * i;
            const int r = q / 8, kc = (q % 8) * 8;
#pragma unroll
            for (int j = 0; j < 8; j++) v[i][j] = t[kc + j][r];
        }
    }
    const int64_t rows = task.which == 0 ? d.mp : d.np;
    E *dst = ops + (task.which == 0 ? d.a_off : d.b_off);
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int q = tid + 256 * i;
        const int64_t row = r0 + q / 8, col = k0 + (q % 8) * 8;
        put8<P>(dst + row * d.kp + col, rows * d.kp, v[i]);
    }
}


// Workgroups of the GEMM that fit a CU's 160 KB of LDS (at most 4), and
// the waves per SIMD that makes: the registers are held to match, so that
// one workgroup's barrier and LDS staging overlap another's MFMAs.
// LDS row pitch in elements: 16-byte fragment reads of BF16 rows conflict
// at a 144-byte pitch and not at 160 bytes (measured); other rows carry
// one 16-byte chunk of padding.
template <int P, int KSTEP> struct GemmPitch {
    static constexpr int value =
        sizeof(typename Elem<P>::T) == 2 && KSTEP == 64 ? KSTEP + 16 : KSTEP + 16 / (int)sizeof(typename Elem<P>::T);
};

template <int P, int BM, int BN, int WAVES, int KSTEP> struct GemmOcc {
    static constexpr int PITCH = GemmPitch<P, KSTEP>::value;
    static constexpr int LDS = 2 * Elem<P>::planes * (BM + BN) * PITCH * (int)sizeof(typename Elem<P>::T);
    static constexpr int groups = 163840 / LDS < 4 ? 163840 / LDS : 4;
    static constexpr int per_simd = groups * WAVES / 4 > 0 ? groups * WAVES / 4 : 1;
};

// One workgroup of WGM x WGN waves computes a BM x BN output tile; wave
// (wr, wc) owns a (BM/WGM) x (BN/WGN) block as 16 x 16 MFMA blocks, k in
// steps of KSTEP.  Larger tiles stage fewer operand bytes per product:
// 256 x 256 does 128 FLOP per staged byte, 128 x 128 64.
template <int P, int BM, int BN, int WGM, int WGN, int KSTEP>
__global__ __launch_bounds__(64 * WGM * WGN)
__attribute__((amdgpu_waves_per_eu(GemmOcc<P, BM, BN, WGM * WGN, KSTEP>::per_simd)))
void grouped_gemm_kernel(const GemmDesc *descs, const Item *items, const int *wg_off, const typename Elem<P>::T *ops) {
    typedef typename Elem<P>::T E;
    typedef typename Acc<P>::type AccV;
    constexpr int NT = 64 * WGM * WGN;
    constexpr int PL = Elem<P>::planes;
    constexpr int CE = 16 / sizeof(E);
    constexpr int PITCH = GemmPitch<P, KSTEP>::value;
    constexpr int APLANE = BM * PITCH, BPLANE = BN * PITCH;
    constexpr int BUF = PL * (APLANE + BPLANE);
    constexpr int ACH = PL * BM * (KSTEP / CE), BCH = PL * BN * (KSTEP / CE);
    constexpr int AN = ACH / NT, BNN = BCH / NT, N = AN + BNN;
    constexpr int WM = BM / WGM, WN = BN / WGN, IM = WM / 16, JN = WN / 16;
    constexpr int STEPS = KSTEP / 32;   // 32-wide fragment chunks per k step
    static_assert(ACH % NT == 0 && BCH % NT == 0 && KSTEP % 32 == 0 && 64 % KSTEP == 0, "tile shape");
    __shared__ __attribute__((aligned(16))) E lds[2 * BUF];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, c = lane & 15;
    const int wr = w / WGN, wc = w % WGN;
    // This thread's chunks: (plane, row, col) of the A and B tiles.
    int crow_[N], ccol[N], cpl[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int kq = tid + NT * (i < AN ? i : i - AN);
        const int per_row = KSTEP / CE;
        const int rows = i < AN ? BM : BN;
        cpl[i] = kq / (rows * per_row);
        crow_[i] = (kq % (rows * per_row)) / per_row;
        ccol[i] = (kq % per_row) * CE;
    }
    // Workgroup g computes items wg_off[g] .. wg_off[g + 1] - 1 (the host
    // balanced the CUs' shares).
    // The next item's first k step is loaded into registers during this
    // item's last step (and its epilogue), so consecutive items' loads
    // overlap: most items here have few k steps.
    u32x4 pre[N];
    const int it_end = wg_off[blockIdx.x + 1];
    auto gload = [&](const E *A_, const E *B_, int64_t aplane_, int64_t bplane_, int64_t kp_, int64_t k0_)
        __attribute__((always_inline)) {
#pragma unroll
        for (int i_ = 0; i_ < N; i_++) {
            const E *src_ = i_ < AN ? A_ + cpl[i_] * aplane_ : B_ + cpl[i_] * bplane_;
            pre[i_] = *reinterpret_cast<const u32x4 *>(src_ + (int64_t)crow_[i_] * kp_ + k0_ + ccol[i_]);
        }
    };
    // Item `it`'s operand bases, plane strides and padded k.
    struct Ops { const E *A, *B; int64_t aplane, bplane, kp; };
    auto ops_of = [&](const Item &item, const GemmDesc &d) __attribute__((always_inline)) {
        Ops o;
        o.A = ops + d.a_off + (int64_t)item.tm * BM * d.kp;
        o.B = ops + d.b_off + (int64_t)item.tn * BN * d.kp;
        o.aplane = d.mp * d.kp;
        o.bplane = d.np * d.kp;
        o.kp = d.kp;
        return o;
    };
    if (wg_off[blockIdx.x] < it_end) {
        const Item first = items[wg_off[blockIdx.x]];
        const Ops o = ops_of(first, descs[first.problem]);
        if (o.kp > 0) gload(o.A, o.B, o.aplane, o.bplane, o.kp, 0);
    }
    for (int it = wg_off[blockIdx.x]; it < it_end; it++) {
        const Item item = items[it];
        const GemmDesc d = descs[item.problem];
        const int64_t m0 = (int64_t)item.tm * BM, n0 = (int64_t)item.tn * BN;
        const Ops o = ops_of(item, d);
        const bool has_next = it + 1 < it_end;
        Ops no{};
        if (has_next) {
            const Item nitem = items[it + 1];
            no = ops_of(nitem, descs[nitem.problem]);
        }
        AccV acc[IM][JN];
#pragma unroll
        for (int i = 0; i < IM; i++)
#pragma unroll
            for (int j = 0; j < JN; j++) acc[i][j] = AccV{0, 0, 0, 0};
#define EF_LSTORE(BUFP)                                                                               \
    _Pragma("unroll") for (int i_ = 0; i_ < N; i_++) {                                               \
        E *dst_ = (BUFP) + (i_ < AN ? cpl[i_] * APLANE : PL * APLANE + cpl[i_] * BPLANE);             \
        *reinterpret_cast<u32x4 *>(dst_ + crow_[i_] * PITCH + ccol[i_]) = pre[i_];                    \
    }
        const int64_t nk = d.kp / KSTEP;   // 0 for k = 0: no loads, C = 0
        if (nk > 0) EF_LSTORE(lds);        // step 0, loaded before this item began
        __syncthreads();
        for (int64_t ks = 0; ks < nk; ks++) {
            const int cur = (int)(ks & 1);
            if (ks + 1 < nk) gload(o.A, o.B, o.aplane, o.bplane, o.kp, (ks + 1) * KSTEP);
            else if (has_next && no.kp > 0) gload(no.A, no.B, no.aplane, no.bplane, no.kp, 0);
            const E *as = lds + cur * BUF;
            const E *bs = as + PL * APLANE;
#pragma unroll
            for (int st = 0; st < STEPS; st++) {
                Frag<P> af[IM], bf[JN];
#pragma unroll
                for (int i = 0; i < IM; i++) tile_load8<P>(af[i], as, APLANE, (wr * WM + i * 16 + c) * PITCH + st * 32 + g * 8);
#pragma unroll
                for (int j = 0; j < JN; j++) tile_load8<P>(bf[j], bs, BPLANE, (wc * WN + j * 16 + c) * PITCH + st * 32 + g * 8);
#pragma unroll
                for (int sp = 0; sp < Parts<P>::n; sp++)
#pragma unroll
                    for (int i = 0; i < IM; i++)
#pragma unroll
                        for (int j = 0; j < JN; j++) mma_part<P>(acc[i][j], af[i], bf[j], sp);
                // All of the chunk's fragment reads before its MFMAs: the
                // reads' latencies overlap instead of one wait per fragment.
                if constexpr (P != PREC_F32) {
                    __builtin_amdgcn_sched_group_barrier(0x100, Elem<P>::planes * (IM + JN), 0);
                    __builtin_amdgcn_sched_group_barrier(0x008, Parts<P>::n * IM * JN, 0);
                }
            }
            if (ks + 1 < nk) EF_LSTORE(lds + (cur ^ 1) * BUF);
            __syncthreads();
        }
        if (nk == 0 && has_next && no.kp > 0) gload(no.A, no.B, no.aplane, no.bplane, no.kp, 0);
#undef EF_LSTORE
        // Lane holds C[m0 + wr*WM + i*16 + crow(g, r)][n0 + wc*WN + j*16 + c].
#pragma unroll
        for (int i = 0; i < IM; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wr * WM + i * 16 + crow<P>(g, r);
                if (row >= d.m) continue;
#pragma unroll
                for (int j = 0; j < JN; j++) {
                    const int64_t col = n0 + wc * WN + j * 16 + c;
                    if (col < d.n) d.c[row * d.n + col] = acc[i][j][r];
                }
            }
    }
}

// ---- operands loaded straight into LDS -----------------------------------

typedef __attribute__((address_space(3))) void lds_void;

template <class T> __device__ __forceinline__ unsigned lds_addr(T *p) { return (unsigned)(uintptr_t)(lds_void *)p; }

// 16 bytes per lane from global memory into LDS at lds + 16 lane.
__device__ __forceinline__ void dma16(const void *g, unsigned lds) {
    __builtin_amdgcn_global_load_lds(g, (lds_void *)(uintptr_t)lds, 16, 0, 0);
}

// PLN planes of R rows x C elements of a prepared operand (row stride
// gstride, plane stride gplane, in elements) into LDS rows of PITCH
// elements, plane after plane, by 1 KiB global_load_lds instructions (lane
// l writes the 16 bytes at 16 l) issued round-robin by the NW waves.
// Padding slots, and slots past the data, load a copy of the first chunk
// of their row.
template <class E, int PLN, int R, int C, int PITCH, int NW> struct DmaTile {
    static constexpr int CE = 16 / (int)sizeof(E), SLOTS = PITCH / CE, DATA = C / CE;
    static constexpr int TOTAL = PLN * R * SLOTS;
    static constexpr int NI = (TOTAL + 63) / 64;
    static constexpr int BYTES = NI * 1024;
    static constexpr int PER = (NI + NW - 1) / NW;   // instructions per wave (n = w + NW i)
    static_assert(PITCH % CE == 0 && C % CE == 0, "rows fill whole 16-byte slots");
    // This lane's source offsets (elements): fixed for a problem.
    __device__ static __forceinline__ void offsets(int64_t (&off)[PER], int64_t gstride, int64_t gplane) {
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int sl = (w + NW * i) * 64 + lane;
            off[i] = 0;
            if (sl < TOTAL) {
                const int pl = sl / (R * SLOTS), rem = sl % (R * SLOTS);
                const int r = rem / SLOTS, col = rem % SLOTS < DATA ? rem % SLOTS : 0;
                off[i] = pl * gplane + (int64_t)r * gstride + col * CE;
            }
        }
    }
    __device__ static __forceinline__ void issue(const E *g, const int64_t (&off)[PER], unsigned lds) {
        const int w = threadIdx.x >> 6;
#pragma unroll
        for (int i = 0; i < PER; i++)
            if (w + NW * i < NI) dma16(g + off[i], lds + (w + NW * i) * 1024);
    }
};

template <int> struct Parity {};

// The GEMM with each k step's A and B tiles loaded straight into one of
// two LDS buffers (separate arrays, the loop unrolled by parity, so that
// reads of one buffer need not wait for the loads into the other): no
// registers stage the next step and no LDS stores.  The operands are
// padded to whole tiles, so every load is in bounds.  Used for BF16 at
// 256 x 256 (9 % faster than staging through registers at 8192^3); the
// BF16x3 tiles, with 32-wide k steps and a fifth of each load on padding
// slots, are faster staged through registers.
template <int P, int BM, int BN, int WGM, int WGN, int KSTEP>
__global__ __launch_bounds__(64 * WGM * WGN)
__attribute__((amdgpu_waves_per_eu(GemmOcc<P, BM, BN, WGM * WGN, KSTEP>::per_simd)))
void grouped_gemm_dma_kernel(const GemmDesc *descs, const Item *items, const int *wg_off, const typename Elem<P>::T *ops) {
    typedef typename Elem<P>::T E;
    typedef typename Acc<P>::type AccV;
    constexpr int NW = WGM * WGN;
    constexpr int PL = Elem<P>::planes;
    constexpr int PITCH = GemmPitch<P, KSTEP>::value;
    constexpr int APLANE = BM * PITCH, BPLANE = BN * PITCH;
    constexpr int WM = BM / WGM, WN = BN / WGN, IM = WM / 16, JN = WN / 16;
    constexpr int STEPS = KSTEP / 32;
    typedef DmaTile<E, PL, BM, KSTEP, PITCH, NW> TA;
    typedef DmaTile<E, PL, BN, KSTEP, PITCH, NW> TB;
    constexpr int BUFB = TA::BYTES + TB::BYTES;
    __shared__ __attribute__((aligned(1024))) char buf0[BUFB];
    __shared__ __attribute__((aligned(1024))) char buf1[BUFB];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, c = lane & 15;
    const int wr = w / WGN, wc = w % WGN;
    for (int it = wg_off[blockIdx.x]; it < wg_off[blockIdx.x + 1]; it++) {
        const Item item = items[it];
        const GemmDesc d = descs[item.problem];
        const int64_t m0 = (int64_t)item.tm * BM, n0 = (int64_t)item.tn * BN;
        const E *A = ops + d.a_off + m0 * d.kp;
        const E *B = ops + d.b_off + n0 * d.kp;
        int64_t aoff[TA::PER], boff[TB::PER];
        TA::offsets(aoff, d.kp, d.mp * d.kp);
        TB::offsets(boff, d.kp, d.np * d.kp);
        AccV acc[IM][JN];
#pragma unroll
        for (int i = 0; i < IM; i++)
#pragma unroll
            for (int j = 0; j < JN; j++) acc[i][j] = AccV{0, 0, 0, 0};
        const int64_t nk = d.kp / KSTEP;   // 0 for k = 0: no loads, C = 0
        auto compute = [&](const char *bufp) {
            const E *as = reinterpret_cast<const E *>(bufp);
            const E *bs = reinterpret_cast<const E *>(bufp + TA::BYTES);
#pragma unroll
            for (int st = 0; st < STEPS; st++) {
                Frag<P> af[IM], bf[JN];
#pragma unroll
                for (int i = 0; i < IM; i++) tile_load8<P>(af[i], as, APLANE, (wr * WM + i * 16 + c) * PITCH + st * 32 + g * 8);
#pragma unroll
                for (int j = 0; j < JN; j++) tile_load8<P>(bf[j], bs, BPLANE, (wc * WN + j * 16 + c) * PITCH + st * 32 + g * 8);
#pragma unroll
                for (int sp = 0; sp < Parts<P>::n; sp++)
#pragma unroll
                    for (int i = 0; i < IM; i++)
#pragma unroll
                        for (int j = 0; j < JN; j++) mma_part<P>(acc[i][j], af[i], bf[j], sp);
                __builtin_amdgcn_sched_group_barrier(0x100, Elem<P>::planes * (IM + JN), 0);
                __builtin_amdgcn_sched_group_barrier(0x008, Parts<P>::n * IM * JN, 0);
            }
        };
        // One k step: the next step's tiles into the other buffer, this
        // step's products, then every load landed and every wave done.
        auto step = [&](auto par, int64_t ks) {
            constexpr bool odd = std::is_same<decltype(par), Parity<1>>::value;
            if (ks + 1 < nk) {
                TA::issue(A + (ks + 1) * KSTEP, aoff, lds_addr(odd ? buf0 : buf1));
                TB::issue(B + (ks + 1) * KSTEP, boff, lds_addr((odd ? buf0 : buf1) + TA::BYTES));
            }
            compute(odd ? buf1 : buf0);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        };
        if (nk > 0) {
            TA::issue(A, aoff, lds_addr(buf0));
            TB::issue(B, boff, lds_addr(buf0 + TA::BYTES));
            __builtin_amdgcn_s_waitcnt(0);
        }
        __syncthreads();
        for (int64_t ks = 0; ks < nk; ks += 2) {
            step(Parity<0>{}, ks);
            if (ks + 1 < nk) step(Parity<1>{}, ks + 1);
        }
        // Lane holds C[m0 + wr*WM + i*16 + crow(g, r)][n0 + wc*WN + j*16 + c].
#pragma unroll
        for (int i = 0; i < IM; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wr * WM + i * 16 + crow<P>(g, r);
                if (row >= d.m) continue;
#pragma unroll
                for (int j = 0; j < JN; j++) {
                    const int64_t col = n0 + wc * WN + j * 16 + c;
                    if (col < d.n) d.c[row * d.n + col] = acc[i][j][r];
                }
            }
    }
}

size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// Tile shapes: config 0 is 128 x 128, 1 is 64 x 64, 2 the large shape
// (256 x 256 for BF16; 256 x 128 with 32-wide k steps for BF16x3, whose two
// planes would not fit LDS twice at 256 x 256, and for FP32, whose
// 128 x 128 per wave at 256 x 256 would spill registers).
bool config_valid(uint32_t precision, int config) { return config <= 2; }
int tile_m(int config) { return config == 0 ? 128 : config == 1 ? 64 : 256; }
int tile_n(uint32_t precision, int config) {
    return config == 0 ? 128 : config == 1 ? 64 : (precision == EF_PRECISION_BF16 ? 256 : 128);
}

}  // namespace

// A prepared set of products: descriptors, prep tasks and work items on the
// device, the prepared operands, and an instantiated graph of
// prep + counter reset + GEMM on the plan's stream.
struct GroupedPlan {
    std::vector<ef_gemm_problem> problems;
    std::vector<double> density;      // per problem: operand rows over cells (1 for device tensors)
    uint32_t precision = EF_PRECISION_BF16X3;
    int config = 0;                   // tile shape: 0 = 128 x 128, 1 = 64 x 64, 2 = large (BF16 precisions)
    DeviceBuffer mem;
    GemmDesc *descs = nullptr;
    PrepTask *tasks = nullptr;
    Item *items = nullptr;
    int *wg_off = nullptr;            // grid + 1: workgroup g's items are wg_off[g] .. wg_off[g + 1] - 1
    void *ops = nullptr;
    int n_tasks = 0, n_items = 0, grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GroupedPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

int cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) ==
                                               hipSuccess && n > 0)
        return n;
    return 256;
}

// Workgroups of the plan's GEMM kernel that one CU holds at once (by the
// runtime's occupancy calculator; 1 if it cannot tell).
template <class K> int occupancy(K kernel, int threads) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        return 1;
    }
    return n;
}

template <int P> int gemm_slots_of(int config) {
    if (config == 0) return occupancy(grouped_gemm_kernel<P, 128, 128, 2, 2, 64>, 256);
    if (config == 1) return occupancy(grouped_gemm_kernel<P, 64, 64, 2, 2, 64>, 256);
    if constexpr (P == PREC_BF16) return occupancy(grouped_gemm_dma_kernel<P, 256, 256, 2, 4, 64>, 512);
    if constexpr (P == PREC_BF16X3) return occupancy(grouped_gemm_kernel<P, 256, 128, 4, 2, 32>, 512);
    if constexpr (P == PREC_F32) return occupancy(grouped_gemm_kernel<P, 256, 128, 2, 2, 32>, 256);
    return 1;
}

int gemm_slots(const GroupedPlan &p) {
    switch (p.precision) {
    case EF_PRECISION_FP32: return gemm_slots_of<PREC_F32>(p.config);
    case EF_PRECISION_BF16X3: return gemm_slots_of<PREC_BF16X3>(p.config);
    case EF_PRECISION_BF16: return gemm_slots_of<PREC_BF16>(p.config);
    default: return 1;
    }
}

ef_status build(GroupedPlan &p, hipStream_t s) {
    const int TM = tile_m(p.config), TN = tile_n(p.precision, p.config);
    const size_t esz = p.precision == EF_PRECISION_FP32 ? 4 : 2;
    const size_t planes = p.precision == EF_PRECISION_BF16X3 ? 2 : 1;
    std::vector<GemmDesc> descs;
    std::vector<PrepTask> tasks;
    std::vector<std::pair<int64_t, Item>> items;
    int64_t off = 0;
    for (size_t i = 0; i < p.problems.size(); i++) {
        const ef_gemm_problem &q = p.problems[i];
        GemmDesc d;
        d.a = q.a; d.b = q.b; d.c = q.c;
        d.m = (int64_t)q.m; d.n = (int64_t)q.n; d.k = (int64_t)q.k;
        // rows padded to whole tiles (of at least 128, the prep's 64 x 64
        // blocks in whole pairs), k to whole 64-wide steps
        const int64_t pm = TM < 128 ? 128 : TM, pn = TN < 128 ? 128 : TN;
        d.mp = (d.m + pm - 1) / pm * pm;
        d.np = (d.n + pn - 1) / pn * pn;
        d.kp = (d.k + KS - 1) / KS * KS;
        d.a_off = off; off += (int64_t)planes * d.mp * d.kp;
        d.b_off = off; off += (int64_t)planes * d.np * d.kp;
        descs.push_back(d);
        if (d.m == 0 || d.n == 0) continue;
        for (int bi = 0; bi < d.mp / 64; bi++)
            for (int bj = 0; bj < d.kp / 64; bj++) tasks.push_back({(int)i, 0, bi, bj});
        for (int bi = 0; bi < d.np / 64; bi++)
            for (int bj = 0; bj < d.kp / 64; bj++) tasks.push_back({(int)i, 1, bi, bj});
        for (int tm = 0; tm < (d.m + TM - 1) / TM; tm++)
            for (int tn = 0; tn < (d.n + TN - 1) / TN; tn++) items.push_back({d.kp, Item{(int)i, tm, tn, 0}});
    }
    // Largest k first; ties in problem and tile order (deterministic list).
    std::stable_sort(items.begin(), items.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
    // Workgroups: as many per CU as fit at once (SLOTS; the hardware puts
    // workgroups g, g + C, g + 2 C, ... on one CU of C), at most one per
    // item.  Each item goes,
    // heaviest first, to the CU with the least k steps so far (with a free
    // workgroup while items fit one per workgroup), then to that CU's
    // least loaded workgroup: every CU gets about the same work, where the
    // round-robin order gave some CUs the heaviest item of every round.
    const int SLOTS = gemm_slots(p);
    const int C = cus();
    const int n = (int)items.size();
    const int grid = n <= C ? n : SLOTS * C;
    std::vector<std::vector<int>> per_wg((size_t)std::max(grid, 1));
    if (n <= C) {
        for (int i = 0; i < n; i++) per_wg[i].push_back(i);
    } else {
        const bool one_each = n <= SLOTS * C;
        std::vector<int64_t> cu_load(C, 0), wg_load(grid, 0);
        std::vector<int> cu_used(C, 0);
        for (int i = 0; i < n; i++) {
            int best = -1;
            for (int c = 0; c < C; c++) {
                if (one_each && cu_used[c] >= SLOTS) continue;
                if (best < 0 || cu_load[c] < cu_load[best]) best = c;
            }
            int wg = -1;
            for (int j = 0; j < SLOTS; j++) {
                const int g = best + j * C;
                if (g >= grid) break;
                if (one_each && !per_wg[g].empty()) continue;
                if (wg < 0 || wg_load[g] < wg_load[wg]) wg = g;
            }
            per_wg[wg].push_back(i);
            cu_load[best] += items[i].first;
            wg_load[wg] += items[i].first;
            cu_used[best]++;
        }
    }
    std::vector<int> wg_off(grid + 1, 0);
    std::vector<std::pair<int64_t, Item>> ordered;
    for (int g = 0; g < grid; g++) {
        wg_off[g] = (int)ordered.size();
        for (int i : per_wg[g]) ordered.push_back(items[i]);
    }
    wg_off[grid] = (int)ordered.size();
    items.swap(ordered);
    p.n_tasks = (int)tasks.size();
    p.n_items = (int)items.size();
    const size_t bytes = up256(descs.size() * sizeof(GemmDesc)) + up256(tasks.size() * sizeof(PrepTask)) +
                         up256(items.size() * sizeof(Item)) + up256(wg_off.size() * sizeof(int)) +
                         up256((size_t)off * esz);
    EF_TRY(p.mem.alloc(bytes));
    char *at = static_cast<char *>(p.mem.ptr);
    p.descs = reinterpret_cast<GemmDesc *>(at); at += up256(descs.size() * sizeof(GemmDesc));
    p.tasks = reinterpret_cast<PrepTask *>(at); at += up256(tasks.size() * sizeof(PrepTask));
    p.items = reinterpret_cast<Item *>(at); at += up256(items.size() * sizeof(Item));
    p.wg_off = reinterpret_cast<int *>(at); at += up256(wg_off.size() * sizeof(int));
    p.ops = at;
    std::vector<Item> flat;
    for (auto &x : items) flat.push_back(x.second);
    if (!descs.empty()) EF_HIP(hipMemcpyAsync(p.descs, descs.data(), descs.size() * sizeof(GemmDesc), hipMemcpyHostToDevice, s));
    if (!tasks.empty()) EF_HIP(hipMemcpyAsync(p.tasks, tasks.data(), tasks.size() * sizeof(PrepTask), hipMemcpyHostToDevice, s));
    if (!flat.empty()) EF_HIP(hipMemcpyAsync(p.items, flat.data(), flat.size() * sizeof(Item), hipMemcpyHostToDevice, s));
    EF_HIP(hipMemcpyAsync(p.wg_off, wg_off.data(), wg_off.size() * sizeof(int), hipMemcpyHostToDevice, s));
    EF_HIP(hipStreamSynchronize(s));
    p.grid = grid;
    return EF_OK;
}

template <int P>
ef_status launch_steps(const GroupedPlan &p, hipStream_t s) {
    typedef typename Elem<P>::T E;
    if (p.n_items == 0) return EF_OK;
    hipLaunchKernelGGL(grouped_prep_kernel<P>, dim3((unsigned)p.n_tasks), dim3(256), 0, s, p.descs, p.tasks, (E *)p.ops);
    if (p.config == 0) {
        hipLaunchKernelGGL((grouped_gemm_kernel<P, 128, 128, 2, 2, 64>), dim3((unsigned)p.grid), dim3(256), 0, s,
                           p.descs, p.items, p.wg_off, (const E *)p.ops);
    } else if (p.config == 1) {
        hipLaunchKernelGGL((grouped_gemm_kernel<P, 64, 64, 2, 2, 64>), dim3((unsigned)p.grid), dim3(256), 0, s,
                           p.descs, p.items, p.wg_off, (const E *)p.ops);
    } else if constexpr (P == PREC_BF16) {
        hipLaunchKernelGGL((grouped_gemm_dma_kernel<P, 256, 256, 2, 4, 64>), dim3((unsigned)p.grid), dim3(512), 0, s,
                           p.descs, p.items, p.wg_off, (const E *)p.ops);
    } else if constexpr (P == PREC_BF16X3) {
        hipLaunchKernelGGL((grouped_gemm_kernel<P, 256, 128, 4, 2, 32>), dim3((unsigned)p.grid), dim3(512), 0, s,
                           p.descs, p.items, p.wg_off, (const E *)p.ops);
    } else {
        // FP32: 256 x 128, four waves of 128 x 64, 32-wide k steps
        hipLaunchKernelGGL((grouped_gemm_kernel<P, 256, 128, 2, 2, 32>), dim3((unsigned)p.grid), dim3(256), 0, s,
                           p.descs, p.items, p.wg_off, (const E *)p.ops);
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status launch_all(const GroupedPlan &p, hipStream_t s) {
    switch (p.precision) {
    case EF_PRECISION_FP32: return launch_steps<PREC_F32>(p, s);
    case EF_PRECISION_BF16X3: return launch_steps<PREC_BF16X3>(p, s);
    case EF_PRECISION_BF16: return launch_steps<PREC_BF16>(p, s);
    default: set_error("grouped products take FP32, BF16X3 or BF16"); return EF_INVALID;
    }
}

// The autotuner's memory: the fastest tile shape per (precision, set of
// (M, N, K, key density)), measured once per process.
std::mutex tune_mu;
std::map<std::vector<int64_t>, int> tuned;

std::vector<int64_t> signature(const GroupedPlan &p) {
    // Each problem's M, N and K bucketed to powers of two and its key
    // density (the fold's operand rows over the cells they fill) to eighths,
    // sorted: similar sets share a choice.
    std::vector<int64_t> sig = {(int64_t)p.precision};
    std::vector<int64_t> shapes;
    auto bucket = [](size_t v) { int64_t b = 0; while (((size_t)1 << b) < v) b++; return b; };
    for (size_t i = 0; i < p.problems.size(); i++) {
        const ef_gemm_problem &q = p.problems[i];
        const double d = i < p.density.size() ? p.density[i] : 1.0;
        const int64_t eighths = d >= 1.0 ? 8 : (int64_t)(d * 8.0);
        shapes.push_back(bucket(q.m) << 24 | bucket(q.n) << 16 | bucket(q.k) << 8 | eighths);
    }
    std::sort(shapes.begin(), shapes.end());
    sig.insert(sig.end(), shapes.begin(), shapes.end());
    return sig;
}

// The tile shape for a plan: the caller's (64, 128 or, for the BF16
// precisions, 256), else the autotuner's choice for its signature, timed
// now if this process has not met the signature before.
ef_status choose_config(GroupedPlan &p, int tile, hipStream_t s);

ef_status time_config(GroupedPlan &p, hipStream_t s, float &ms) {
    EF_TRY(build(p, s));
    hipEvent_t a, b;
    EF_HIP(hipEventCreate(&a));
    EF_HIP(hipEventCreate(&b));
    EF_TRY(launch_all(p, s));   // warm
    EF_HIP(hipEventRecord(a, s));
    for (int r = 0; r < 3; r++) EF_TRY(launch_all(p, s));
    EF_HIP(hipEventRecord(b, s));
    EF_HIP(hipEventSynchronize(b));
    EF_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return EF_OK;
}

}  // namespace

namespace {

ef_status choose_config(GroupedPlan &p, int tile, hipStream_t s) {
    if (tile == 128 || tile == 64 || (tile == 256 && config_valid(p.precision, 2))) {
        p.config = tile == 128 ? 0 : tile == 64 ? 1 : 2;
        return EF_OK;
    }
    const auto sig = signature(p);
    {
        std::lock_guard<std::mutex> lock(tune_mu);
        auto it = tuned.find(sig);
        if (it != tuned.end()) {
            p.config = it->second;
            return EF_OK;
        }
    }
    float best = 1e30f;
    int best_cfg = 0;
    for (int cfg = 0; cfg < 3; cfg++) {
        if (!config_valid(p.precision, cfg)) continue;
        p.config = cfg;
        float ms = 0;
        EF_TRY(time_config(p, s, ms));
        if (ms < best) { best = ms; best_cfg = cfg; }
    }
    p.config = best_cfg;
    std::lock_guard<std::mutex> lock(tune_mu);
    tuned[sig] = best_cfg;
    return EF_OK;
}

}  // namespace

ef_status grouped_run_once(const ef_gemm_problem *problems, size_t count, uint32_t precision, const double *density,
                           hipStream_t s) {
    if (precision != EF_PRECISION_FP32 && precision != EF_PRECISION_BF16X3 && precision != EF_PRECISION_BF16) {
        set_error("grouped products take FP32, BF16X3 or BF16");
        return EF_INVALID;
    }
    GroupedPlan plan;
    plan.problems.assign(problems, problems + count);
    if (density) plan.density.assign(density, density + count);
    plan.precision = precision;
    EF_TRY(choose_config(plan, 0, s));
    EF_TRY(build(plan, s));
    EF_TRY(launch_all(plan, s));
    // The plan's buffers are freed on return: wait for the launches.
    EF_HIP(hipStreamSynchronize(s));
    return EF_OK;
}

ef_status grouped_plan_create(const ef_gemm_problem *problems, size_t count, uint32_t precision, int tile,
                              void *stream, GroupedPlan **out) {
    *out = nullptr;
    hipStream_t s = as_stream(stream);
    if (precision != EF_PRECISION_FP32 && precision != EF_PRECISION_BF16X3 && precision != EF_PRECISION_BF16) {
        set_error("grouped products take FP32, BF16X3 or BF16");
        return EF_INVALID;
    }
    auto plan = new GroupedPlan();
    plan->problems.assign(problems, problems + count);
    plan->precision = precision;
    ef_status st = choose_config(*plan, tile, s);
    if (st == EF_OK) st = build(*plan, s);
    // One graph of the three launches, replayed by each run.
    if (st == EF_OK && plan->n_items > 0) {
        hipStream_t cap;
        if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
            st = hip_status(hipGetLastError(), "hipStreamCreate");
        } else {
            hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                st = launch_all(*plan, cap);
                hipGraph_t g = nullptr;
                hipError_t e2 = hipStreamEndCapture(cap, &g);
                if (st == EF_OK && e2 == hipSuccess) {
                    plan->graph = g;
                    e2 = hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0);
                }
                if (st == EF_OK && e2 != hipSuccess) st = hip_status(e2, "graph capture");
            } else {
                st = hip_status(e, "hipStreamBeginCapture");
            }
            (void)hipStreamDestroy(cap);
        }
    }
    if (st != EF_OK) {
        delete plan;
        return st;
    }
    *out = plan;
    return EF_OK;
}

ef_status grouped_plan_run(GroupedPlan *plan, void *stream) {
    if (!plan) {
        set_error("no plan");
        return EF_INVALID;
    }
    if (plan->n_items == 0) return EF_OK;
    EF_HIP(hipGraphLaunch(plan->exec, as_stream(stream)));
    return EF_OK;
}

int grouped_plan_tile(const GroupedPlan *plan) { return plan ? tile_m(plan->config) : 0; }

void grouped_plan_destroy(GroupedPlan *plan) { delete plan; }

}  // namespace ef
