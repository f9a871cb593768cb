// Fused attention pipeline on gfx950: the reference's tile algorithm
// (crates/einfold-kernels/src/fused.rs:694-1019) with the score tile in
// registers, an online softmax, and key tiles that no query of a
// workgroup may see never visited; plus backward kernels that recompute the
// score tiles from the saved row max and row sum instead of storing them.
//
// Layout: q [B][Tq][D], k [B][Tk][D], v [B][Tk][D], out [B][Tq][D], row
// statistics [B][Tq].  A prep pass first writes the streamed operands once
// in the form the MFMA fragments read (split into BF16 hi/lo planes for
// BF16x3, transposed where a fragment reads along the other axis), padded
// to whole tiles with zeros, so each tile is staged into LDS by plain
// 16-byte copies.
//
// Scheduling: workgroups are persistent; each takes (batch, block, part)
// items from a counter, the heaviest first.  Under a causal or other
// varying mask a block's range of streamed tiles is cut in two parts,
// each computed by one workgroup into a partial result, and a small
// kernel combines the two in a fixed order: the triangle's long rows are
// then no longer than twice the short ones' work, so no CU idles while
// one finishes them, and the bits never depend on who took which part.  Each item is computed whole by
// one workgroup, keys in order, so the bits never depend on which
// workgroup took it.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "common.h"
#include "mfma.h"

// Compiled a second time as attention_bwd.hip with EF_ATTN_BWD_TU set: that
// unit holds the backward kernels' instances (built with other code
// generation options) and the launcher below, this one everything else.
#ifndef EF_ATTN_BWD_TU
#define EF_ATTN_BWD_TU 0
#endif

namespace ef {

// Launches the dK/dV (kernel 0) or dQ (kernel 1) backward kernel for
// precision P and head size D with the BwdArgs at args.
void attn_bwd_launch(int kernel, int precision, int d, const void *args, unsigned grid, hipStream_t s);

namespace {

constexpr int BN = 64;          // rows of a streamed tile (keys in fwd/dq, queries in dk/dv)
constexpr int WAVES = 4;        // waves per workgroup, 32 owned rows each
constexpr int THREADS = 64 * WAVES;
constexpr int BW = 32 * WAVES;  // rows a workgroup owns
constexpr int PAD = 128;        // prepared operands are padded to whole 128-row blocks
// Row pitch padding (elements) of LDS tiles read by 16-byte fragment
// loads, one row per lane: 32 bytes for BF16 rows (a 16-byte pad leaves
// those loads two-way bank-conflicted on gfx950), 8 elements otherwise.
template <class E> constexpr int RPAD = sizeof(E) == 2 ? 16 : 8;
constexpr double LOG2E = 1.4426950408889634;   // cast to the precision's scalar
constexpr double LN2 = 0.6931471805599453;
constexpr double TAU = 8.0;     // forward weights stay below 2^TAU (log2 units)

struct Span { int64_t lo, hi; };

// Per-row statistics (row max, row sum, delta): float for the FP32 and
// BF16 precisions, double for FP64, so that the FP64 backward recomputes
// its weights from FP64 statistics.
template <int P> struct Stat { typedef float T; };
template <> struct Stat<PREC_F64> { typedef double T; };
__device__ __forceinline__ float log2_of(float x) { return __log2f(x); }
__device__ __forceinline__ double log2_of(double x) { return log2(x); }
struct Span32 { int lo, hi; };

struct Mask {
    const int64_t *allowed;
    int64_t causal, offset, window;
    int64_t tq, tk;
};

// The keys query x may see, [lo, hi), clamped to the keys (fused.rs:350-368).
__device__ __forceinline__ Span query_span(const Mask &m, int64_t x) {
    if (x < 0 || x >= m.tq) return {0, 0};
    int64_t lo = 0, hi = m.tk;
    if (m.allowed) {
        lo = m.allowed[2 * x];
        hi = m.allowed[2 * x + 1];
    } else if (m.causal) {
        int64_t bound = x + m.offset;            // key <= x + offset
        hi = bound + 1 < hi ? bound + 1 : hi;
        if (m.window > 0) {
            int64_t l = bound - m.window + 1;    // key > x + offset - window
            lo = l > lo ? l : lo;
        }
    }
    lo = lo < 0 ? 0 : (lo > m.tk ? m.tk : lo);
    hi = hi < 0 ? 0 : (hi > m.tk ? m.tk : hi);
    if (hi < lo) hi = lo;
    return {lo, hi};
}

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));   // one 16-byte chunk

// Four consecutive values (idx a multiple of 4, the planes 16-byte
// aligned) with one store per plane.
template <int P, class T>
__device__ __forceinline__ void put_split4(typename Elem<P>::T *base, int64_t plane, int64_t idx, const T (&x)[4]) {
    if constexpr (P == PREC_BF16X3) {
        bf16x4 h, l;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            __bf16 hj, lj;
            split_bf16((float)x[j], hj, lj);
            h[j] = hj;
            l[j] = lj;
        }
        *reinterpret_cast<bf16x4 *>(base + idx) = h;
        *reinterpret_cast<bf16x4 *>(base + plane + idx) = l;
    } else if constexpr (P == PREC_BF16) {
        bf16x4 h;
#pragma unroll
        for (int j = 0; j < 4; j++) h[j] = (__bf16)(float)x[j];
        *reinterpret_cast<bf16x4 *>(base + idx) = h;
    } else if constexpr (P == PREC_F32) {
        *reinterpret_cast<f32x4 *>(base + idx) = f32x4{(float)x[0], (float)x[1], (float)x[2], (float)x[3]};
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) base[idx + j] = (double)x[j];
    }
}

// ---------------------------------------------------------------------
// Prep: x [B][T][D] into row-major planes [planes][B][Tp][D] and/or
// transposed planes [planes][B][D][Tp], zero past T.  One block per 64
// rows of one batch.
// ---------------------------------------------------------------------
struct Counter { unsigned next, done; };
constexpr int QUEUES = 8;   // work counters per kernel (one per XCD residue)

// One prep launch: blockIdx.z picks a job.  Job z < njobs prepares one
// operand; job z == njobs (when out is set) computes delta[b][x] = sum_d
// dout * out for 64 rows per block, a fixed-order tree over 16 lanes per
// row.  Block (0, 0, 0) also zeroes the work counters the launch's
// kernels will take items from.
struct PrepJob {
    const void *x;
    int64_t T, Tp;
    void *rows, *trans;
};
struct PrepArgs {
    PrepJob job[4];
    int njobs, nzero;
    int64_t B;
    Counter *zero;
    const void *out, *dout;
    void *delta;                  // Stat<P>::T per query
    int64_t tq;
};

template <int P, int D>
__global__ __launch_bounds__(256) void prep_kernel(PrepArgs a) {
    typedef typename Input<P>::T In;
    typedef typename Elem<P>::T E;
    __shared__ In t[64][D + 1];
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && (int)threadIdx.x < a.nzero) {
        a.zero[threadIdx.x].next = 0;
        a.zero[threadIdx.x].done = 0;
    }
    const int64_t b = blockIdx.y;
    const int64_t t0 = (int64_t)blockIdx.x * 64;
    if ((int)blockIdx.z == a.njobs) {
        // delta rows t0 .. t0 + 63 of batch b
        typedef typename Stat<P>::T St;
        const In *out = static_cast<const In *>(a.out), *dout = static_cast<const In *>(a.dout);
        const int c = threadIdx.x & 15;
        for (int it = 0; it < 4; it++) {
            const int64_t x = t0 + it * 16 + (threadIdx.x >> 4);
            const int64_t r = b * a.tq + x;
            St sm = 0;
            if (x < a.tq)
                for (int e = c; e < D; e += 16) sm += (St)out[r * D + e] * (St)dout[r * D + e];
            for (int off = 8; off >= 1; off >>= 1) sm += __shfl_xor(sm, off, 16);
            if (x < a.tq && c == 0) static_cast<St *>(a.delta)[r] = sm;
        }
        return;
    }
    const PrepJob &j = a.job[blockIdx.z];
    if (t0 >= j.Tp) return;
    const In *x = static_cast<const In *>(j.x);
    E *rows = static_cast<E *>(j.rows), *trans = static_cast<E *>(j.trans);
    const int64_t T = j.T, Tp = j.Tp;
    const int64_t plane = a.B * Tp * D;
    // Four consecutive columns of a row per thread, then four consecutive
    // rows of a column: a store of 8 or 16 bytes per plane instead of 2 or 4.
    for (int i = threadIdx.x; i < 64 * D / 4; i += 256) {
        const int r = i / (D / 4), c = (i % (D / 4)) * 4;
        const int64_t tt = t0 + r;
        In v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            v[j] = tt < T ? x[(b * T + tt) * D + c + j] : In(0);
            t[r][c + j] = v[j];
        }
        if (rows) put_split4<P>(rows, plane, (b * Tp + tt) * D + c, v);
    }
    if (!trans) return;
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * D / 4; i += 256) {
        const int c = i / 16, r = (i % 16) * 4;
        const In v[4] = {t[r][c], t[r + 1][c], t[r + 2][c], t[r + 3][c]};
        put_split4<P>(trans, plane, (b * D + c) * Tp + t0 + r, v);
    }
}

// ---------------------------------------------------------------------
// Work counters: one per kernel, zeroed by a small kernel before each
// launch (plain stores, so the pair may be captured in a graph).
// ---------------------------------------------------------------------
__global__ void zero_counters_kernel(Counter *c) {
    if (threadIdx.x < QUEUES) { c[threadIdx.x].next = 0; c[threadIdx.x].done = 0; }
}

// A workgroup's items: its first is blockIdx.x, the rest (when the grid
// is smaller than the item count) come from a counter, offset by the grid.
// A grid of one workgroup per item never touches the counters: the
// hardware hands out workgroups in index order, heaviest items first.
//
// Items are numbered batch fastest (item % batches is the batch), and the
// hardware deals workgroups to the eight XCDs in turn (workgroup w to XCD
// w % 8), so with a multiple of eight batches each XCD's L2 holds the
// streamed tiles of its own batches only.  Later items keep that: a grid
// that is a multiple of eight takes them from one counter per residue of
// the item number mod 8 (QUEUES counters), item grid + r + 8 n for the
// n-th taken by the workgroups w = r mod 8.  Every item is taken once
// whichever XCD a workgroup really runs on: the residues only pick
// which items share an L2.
struct ItemQueue {
    Counter *ctr;
    int *slot;
    int items;
    __device__ __forceinline__ void start(Counter *c, int *s, int n) {
        ctr = c;
        slot = s;
        items = n;
        if (threadIdx.x == 0) *slot = (int)blockIdx.x;
    }
    __device__ __forceinline__ int next() {
        __syncthreads();
        const int item = *slot;
        __syncthreads();   // all have read slot before thread 0 writes it again
        if (threadIdx.x == 0) {
            const int grid = (int)gridDim.x;
            if (grid >= items) {
                *slot = items;
            } else {
                const int stride = grid % QUEUES == 0 ? QUEUES : 1;
                const int r = (int)blockIdx.x % stride;
                const long long nxt = (long long)grid + r + (long long)stride * atomicAdd(&ctr[r].next, 1u);
                *slot = nxt < items ? (int)nxt : items;
            }
        }
        return item;
    }
};

// Copies `rows` rows of `cols` elements (16-byte chunks) of every plane
// from global (row stride gstride, plane stride gplane) into an LDS tile
// (row pitch lpitch, plane stride lplane).
template <class E>
__device__ __forceinline__ void stage(E *lds, int lpitch, int lplane, const E *g, int64_t gstride, int64_t gplane,
                                      int rows, int cols, int planes) {
    constexpr int CE = 16 / sizeof(E);
    const int per_row = cols / CE;
    const int per_plane = rows * per_row;
    const int total = planes * per_plane;
    for (int k = threadIdx.x; k < total; k += THREADS) {
        const int p = k / per_plane, rem = k % per_plane;
        const int r = rem / per_row, c = (rem % per_row) * CE;
        const u32x4 v = *reinterpret_cast<const u32x4 *>(g + p * gplane + r * gstride + c);
        *reinterpret_cast<u32x4 *>(lds + p * lplane + r * lpitch + c) = v;
    }
}

// A tile of PLN planes x R rows x C elements in 16-byte chunks, spread
// over the workgroup: thread t moves chunks t, t + THREADS, ...  All of a
// thread's loads are issued before any is used, so a tile costs one
// memory latency, and a caller can hold them in registers (prefetch)
// while it computes on the previous tile.
template <class E, int PLN, int R, int C> struct TileIO {
    static constexpr int CE = 16 / sizeof(E);
    static constexpr int PER_ROW = C / CE, PER_PLANE = R * PER_ROW, TOTAL = PLN * PER_PLANE;
    static constexpr int N = (TOTAL + THREADS - 1) / THREADS;
    // Chunk i's source splits into a part common to the workgroup (plane
    // and row block of chunk i) and this thread's 32-bit offset, so that
    // each load is a scalar base plus one register.
    static constexpr bool SPLIT = PER_PLANE % THREADS == 0 && THREADS % PER_ROW == 0;
    // Chunk i alone (a caller may spread a tile's loads over its work).
    __device__ static __forceinline__ void load_chunk(u32x4 (&buf)[N], int i, const E *g, int64_t gstride,
                                                      int64_t gplane) {
        if constexpr (SPLIT) {
            const unsigned lane_off = (unsigned)(((int)threadIdx.x / PER_ROW) * gstride +
                                                 ((int)threadIdx.x % PER_ROW) * CE) * (unsigned)sizeof(E);
            constexpr int BLOCKS = PER_PLANE / THREADS;
            const int p = i / BLOCKS, rows = (i % BLOCKS) * (THREADS / PER_ROW);
            const char *base = reinterpret_cast<const char *>(g + p * gplane + rows * gstride);
            buf[i] = *reinterpret_cast<const u32x4 *>(base + (uint64_t)lane_off);
        } else {
            const int k = (int)threadIdx.x + THREADS * i;
            if (TOTAL % THREADS == 0 || k < TOTAL) {
                const int p = k / PER_PLANE, rem = k % PER_PLANE;
                const int r = rem / PER_ROW, cc = (rem % PER_ROW) * CE;
                buf[i] = *reinterpret_cast<const u32x4 *>(g + p * gplane + r * gstride + cc);
            }
        }
    }
    __device__ static __forceinline__ void load(u32x4 (&buf)[N], const E *g, int64_t gstride, int64_t gplane) {
#pragma unroll
        for (int i = 0; i < N; i++) load_chunk(buf, i, g, gstride, gplane);
    }
    __device__ static __forceinline__ void store(E *lds, int lpitch, int lplane, const u32x4 (&buf)[N]) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            const int k = (int)threadIdx.x + THREADS * i;
            if (TOTAL % THREADS == 0 || k < TOTAL) {
                const int p = k / PER_PLANE, rem = k % PER_PLANE;
                const int r = rem / PER_ROW, cc = (rem % PER_ROW) * CE;
                *reinterpret_cast<u32x4 *>(lds + p * lplane + r * lpitch + cc) = buf[i];
            }
        }
    }
};

// The union [first, last) of the spans of queries q0 .. q0 + BW - 1,
// computed by each wave on its own (BW / 64 queries per lane, a wave-wide
// min / max): no LDS and no barrier.  Empty: first > last.
__device__ __forceinline__ void block_range(const Mask &m, int64_t q0, int64_t &first, int64_t &last) {
    const int lane = threadIdx.x & 63;
    int64_t lo = INT64_MAX, hi = 0;
#pragma unroll
    for (int h = 0; h < BW / 64; h++) {
        const Span sp = query_span(m, q0 + h * 64 + lane);
        if (sp.lo < sp.hi) {
            lo = sp.lo < lo ? sp.lo : lo;
            hi = sp.hi > hi ? sp.hi : hi;
        }
    }
    // Spans lie in [0, tk], tk < 2^31: reduce in 32 bits.
    int l32 = lo == INT64_MAX ? INT32_MAX : (int)lo, h32 = (int)hi;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int l2 = __shfl_xor(l32, o), h2 = __shfl_xor(h32, o);
        l32 = l2 < l32 ? l2 : l32;
        h32 = h2 > h32 ? h2 : h32;
    }
    first = l32 == INT32_MAX ? INT64_MAX : l32;
    last = h32;
}

// ---------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------
struct FwdArgs {
    const void *q;
    const void *kr, *vt;          // prepared K rows and V^T
    void *out;
    void *row_max, *row_sum;      // Stat<P>::T per query
    unsigned long long *visited;
    Mask mask;
    int64_t tq, tk, tkp, batches;
    double scale;
    Counter *ctr;
    int items, reverse, parts, pad;
    void *po, *pm, *pl;           // parts == 2: each part's unnormalized sums, max and sum
};

// The K and V^T tiles of one 64-key step in LDS: K rows (KP elements,
// data then padding) then V^T rows (VP), plane after plane, filled by 1 KiB
// global_load_lds_dwordx4 instructions (lane l writes the 16 bytes at
// 16 l; a padding slot gets a copy of its row's first chunk, slots past
// the tile a copy of the tile's first chunk).  Instruction n is issued by
// wave n % WAVES.
template <int P, int D> struct FwdTiles {
    typedef typename Elem<P>::T E;
    static constexpr int PL = Elem<P>::planes;
    static constexpr int CE = 16 / sizeof(E);
    static constexpr int KP = D + RPAD<E>, VP = BN + 8;
    static constexpr int KPLANE = BN * KP, VPLANE = D * VP;
    static constexpr int KS = KP / CE, VS = VP / CE;                 // 16-byte slots per LDS row
    static constexpr int KSLOTS = PL * BN * KS, VSLOTS = PL * D * VS;
    static constexpr int NK = KSLOTS / 64, NV = (VSLOTS + 63) / 64, N = NK + NV;
    static constexpr int PER = (N + WAVES - 1) / WAVES;              // instructions per wave
    static constexpr int BUF = N * 1024 / (int)sizeof(E);            // elements per LDS buffer
    static constexpr int OCC = (D <= 64 && (P == PREC_BF16X3 || P == PREC_BF16)) ? 2 : 1;   // workgroups per CU
    static_assert(KSLOTS % 64 == 0 && KP % CE == 0 && VP % CE == 0, "tile rows fill whole slots");
    static_assert(OCC == 1 || 2 * BUF * (int)sizeof(E) <= 80 * 1024, "two workgroups' buffers fit in LDS");
};

template <int P, int D>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(FwdTiles<P, D>::OCC)))
void attn_fwd_kernel(FwdArgs a) {
    typedef typename Input<P>::T In;
    typedef typename Acc<P>::type AccV;
    typedef typename Acc<P>::scalar S;
    typedef typename Elem<P>::T E;
    typedef FwdTiles<P, D> F;
    constexpr int KC = D / 32, DB = D / 16, NP = Parts<P>::n;
    __shared__ __attribute__((aligned(1024))) E lds[2 * F::BUF];
    __shared__ int flags[2][WAVES];   // per buffer, per wave: bit 0 any, bit 1 full
    __shared__ int slot;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, c = lane & 15;
    const int64_t nqt = (a.tq + BW - 1) / BW;
    const int64_t kplane = a.batches * a.tkp * D, vplane = a.batches * D * a.tkp;
    const S scale2 = (S)a.scale * (S)LOG2E;   // scores in log2 units
    unsigned long long visited = 0;

    // This lane's source for each of its wave's load instructions, in
    // elements from the K tile's base (n < NK) or the V^T tile's.
    int goff[F::PER];   // fits: a plane of prepared operands stays below 2^31 elements (checked on the host)
#pragma unroll
    for (int i = 0; i < F::PER; i++) {
        const int n = w + WAVES * i;
        goff[i] = 0;
        if (n < F::NK) {
            const int sl = n * 64 + lane, pl = sl / (BN * F::KS), rem = sl % (BN * F::KS);
            const int r = rem / F::KS, col = rem % F::KS < D / F::CE ? rem % F::KS : 0;
            goff[i] = (int)(pl * kplane + (int64_t)r * D + col * F::CE);
        } else if (n < F::N) {
            const int sl = (n - F::NK) * 64 + lane;
            if (sl < F::VSLOTS) {
                const int pl = sl / (D * F::VS), rem = sl % (D * F::VS);
                const int r = rem / F::VS, col = rem % F::VS < BN / F::CE ? rem % F::VS : 0;
                goff[i] = (int)(pl * vplane + (int64_t)r * a.tkp + col * F::CE);
            }
        }
    }
    ItemQueue queue;
    queue.start(a.ctr, &slot, a.items);
    for (;;) {
        const int item = queue.next();
        if (item >= a.items) break;
        const int64_t b = item % a.batches;
        const int64_t rest = item / a.batches;
        const int part = (int)(rest % a.parts);
        const int64_t qt = a.reverse ? nqt - 1 - rest / a.parts : rest / a.parts;
        const int64_t q0 = qt * BW;
        const In *Q = static_cast<const In *>(a.q) + b * a.tq * D;
        const E *K = static_cast<const E *>(a.kr) + b * a.tkp * D;
        const E *VT = static_cast<const E *>(a.vt) + b * D * a.tkp;

        // Query rows and spans in 32 bits (the host keeps tq, tk < 2^31).
        int x[2];
        int64_t xr[2];
        Span32 sp[2];
        Frag<P> qf[2][KC];
#pragma unroll
        for (int qb = 0; qb < 2; qb++) {
            x[qb] = (int)q0 + w * 32 + qb * 16 + c;
            // Rows past tq load row tq - 1 (and use 0): unconditional loads
            // issue together instead of one branch and wait each.
            xr[qb] = x[qb] < a.tq ? x[qb] : a.tq - 1;
            const Span s64 = query_span(a.mask, x[qb]);
            sp[qb] = {(int)s64.lo, (int)s64.hi};
#pragma unroll
            for (int kc = 0; kc < KC; kc++) {
                In tmp[8];
#pragma unroll
                for (int j = 0; j < 8; j++) tmp[j] = Q[xr[qb] * D + kc * 32 + g * 8 + j];
#pragma unroll
                for (int j = 0; j < 8; j++) tmp[j] = x[qb] < a.tq ? (In)(tmp[j] * scale2) : In(0);
                make_frag<P>(qf[qb][kc], tmp);
            }
        }
        int64_t first, last;
        block_range(a.mask, q0, first, last);
        int64_t jbeg = first < last ? first - first % BN : last;
        if (a.parts == 2) {
            // Part 0 takes the first half of the key tiles, part 1 the rest.
            const int64_t cut = jbeg + ((last - jbeg + 2 * BN - 1) / (2 * BN)) * BN;
            if (part == 0) last = cut < last ? cut : last;
            else jbeg = cut < last ? cut : last;
        }

        S m[2], mt[2], l[2];
        AccV o[2][DB];
#pragma unroll
        for (int qb = 0; qb < 2; qb++) {
            m[qb] = -INFINITY;
            mt[qb] = -INFINITY;
            l[qb] = 0;
#pragma unroll
            for (int i = 0; i < DB; i++) o[qb][i] = AccV{0, 0, 0, 0};
        }
// Each wave's flags for the tile at J0: bit 0 some query sees it, bit 1
// every query sees all of it.
#define EF_WAVE_FLAGS(J0)                                                                        \
    ([&]() {                                                                                     \
        bool any_ = false, full_ = true;                                                         \
        for (int qb_ = 0; qb_ < 2; qb_++) {                                                      \
            any_ |= sp[qb_].lo < sp[qb_].hi && sp[qb_].lo < (J0) + BN && sp[qb_].hi > (J0);      \
            full_ &= x[qb_] >= a.tq || (sp[qb_].lo <= (J0) && sp[qb_].hi >= (J0) + BN);         \
        }                                                                                        \
        return (__any(any_) ? 1 : 0) | (__all(full_) ? 2 : 0);                                   \
    }())
// Issues this wave's loads of the tile at J0 into LDS buffer BUF.
#define EF_LOAD_TILE(J0, BUF)                                                                    \
    _Pragma("unroll") for (int i_ = 0; i_ < F::PER; i_++) {                                      \
        const int n_ = w + WAVES * i_;                                                           \
        if (n_ < F::N)                                                                           \
            __builtin_amdgcn_global_load_lds(                                                    \
                (const void *)((n_ < F::NK ? K + (J0) * D : VT + (J0)) + goff[i_]),              \
                (__attribute__((address_space(3))) void *)((char *)(BUF) + n_ * 1024), 16, 0, 0);     \
    }
        if (jbeg < last) {
            EF_LOAD_TILE(jbeg, lds);
            const int f0 = EF_WAVE_FLAGS(jbeg);
            if (lane == 0) flags[0][w] = f0;
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        int cur = 0;
        for (int64_t j0 = jbeg; j0 < last; j0 += BN) {
            const bool more = j0 + BN < last;
            if (more) {
                // The next tile streams into the other buffer meanwhile.
                EF_LOAD_TILE(j0 + BN, lds + (cur ^ 1) * F::BUF);
                const int fn = EF_WAVE_FLAGS(j0 + BN);
                if (lane == 0) flags[cur ^ 1][w] = fn;
            }
            int f_and = 3;
#pragma unroll
            for (int i = 0; i < WAVES; i++) f_and &= flags[cur][i];
            const bool full_wg = (f_and & 2) != 0;
            // Tiles in the reference's 64 x 64 units: each 64 queries (two
            // waves) of the workgroup count the tiles they see.
            if (tid == 0)
#pragma unroll
                for (int h = 0; h < WAVES / 2; h++) visited += (unsigned)(((flags[cur][2 * h] | flags[cur][2 * h + 1]) & 1) != 0);
            const E *ks = lds + cur * F::BUF;
            const E *vt = ks + F::PL * F::KPLANE;

            // FP32 accumulators: the scores arrive less the reference max
            // (they start at -base, base = m, or 0 before any key is seen),
            // so the weights need no subtraction unless the reference moves.
            // FP64 (base 0) keeps the subtraction: its registers are full.
            constexpr bool FOLD = P != PREC_F64;
            S base[2];
            AccV s[2][4];
#pragma unroll
            for (int qb = 0; qb < 2; qb++) {
                base[qb] = FOLD && m[qb] != -INFINITY ? m[qb] : S(0);
#pragma unroll
                for (int kb = 0; kb < 4; kb++) s[qb][kb] = AccV{-base[qb], -base[qb], -base[qb], -base[qb]};
            }
#pragma unroll
            for (int kc = 0; kc < KC; kc++) {
#pragma unroll
                for (int kh = 0; kh < 4; kh += 2) {   // two key blocks' fragments live at a time
                    Frag<P> kf[2];
#pragma unroll
                    for (int kb = 0; kb < 2; kb++)
                        tile_load8<P>(kf[kb], ks, F::KPLANE, ((kh + kb) * 16 + c) * F::KP + kc * 32 + g * 8);
#pragma unroll
                    for (int sp_ = 0; sp_ < NP; sp_++)
#pragma unroll
                        for (int kb = 0; kb < 2; kb++)
#pragma unroll
                            for (int qb = 0; qb < 2; qb++) mma_part<P>(s[qb][kh + kb], kf[kb], qf[qb][kc], sp_);
                }
            }
            // Online softmax per query column (the query fragments carry
            // scale * log2(e), so the scores are in log2 units), the weights
            // written over the scores.  Branch-free, both query blocks
            // together: a column that has seen no key yet (max -inf) gets
            // weights exp2(-inf - 0) = 0 and keeps l = 0.
            // Masked tiles only: keys outside a query's span get -inf.
            S mk[2][4];
            if (full_wg) {
#pragma unroll
                for (int qb = 0; qb < 2; qb++)
#pragma unroll
                    for (int kb = 0; kb < 4; kb++)
                        mk[qb][kb] = fmax(fmax(s[qb][kb][0], s[qb][kb][1]), fmax(s[qb][kb][2], s[qb][kb][3]));
            } else {
#pragma unroll
                for (int qb = 0; qb < 2; qb++) {
                    const int lo = (int)(sp[qb].lo - j0 < 0 ? -1 : (sp[qb].lo - j0 > BN ? BN : sp[qb].lo - j0));
                    const int hi = (int)(sp[qb].hi - j0 < 0 ? 0 : (sp[qb].hi - j0 > BN ? BN : sp[qb].hi - j0));
#pragma unroll
                    for (int kb = 0; kb < 4; kb++) {
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const int j = kb * 16 + crow<P>(g, i);
                            if (!(j >= lo && j < hi)) s[qb][kb][i] = S(-INFINITY);
                        }
                        mk[qb][kb] = fmax(fmax(s[qb][kb][0], s[qb][kb][1]), fmax(s[qb][kb][2], s[qb][kb][3]));
                    }
                }
            }
            // The weights are taken against a reference max m that moves
            // only when a column's tile max passes it by more than TAU (all
            // weights then stay below 2^TAU); mt keeps the true max.  When
            // no column of the wave moves, o and l need no rescaling (and,
            // with FOLD, the scores are the weights' exponents as they are).
            // tmax is relative to base; m - base is 0, or -inf before any key
            // (FOLD), else m.
            S tmax[2];
            bool move = false;
#pragma unroll
            for (int qb = 0; qb < 2; qb++) {
                tmax[qb] = wave_max(fmax(fmax(mk[qb][0], mk[qb][1]), fmax(mk[qb][2], mk[qb][3])));
                mt[qb] = fmax(mt[qb], tmax[qb] + base[qb]);
                move |= tmax[qb] > (m[qb] - base[qb]) + S(TAU);
            }
            if (__any(move)) {
#pragma unroll
                for (int qb = 0; qb < 2; qb++) {
                    const S mnew = tmax[qb] > (m[qb] - base[qb]) + S(TAU) ? tmax[qb] + base[qb] : m[qb];
                    const S mnew_sub = mnew == -INFINITY ? S(0) : mnew;
                    // 1 when unmoved, 0 when m was -inf (also when it stays so)
                    const S alpha = fast_exp2(m[qb] - mnew_sub);
                    m[qb] = mnew;
                    l[qb] *= alpha;
#pragma unroll
                    for (int db = 0; db < DB; db++)
#pragma unroll
                        for (int i = 0; i < 4; i++) o[qb][db][i] *= alpha;
                    if constexpr (FOLD) {
                        // the exponents against the new reference (0 if unmoved)
                        const S sh = mnew_sub - base[qb];
#pragma unroll
                        for (int kb = 0; kb < 4; kb++)
#pragma unroll
                            for (int i = 0; i < 4; i++) s[qb][kb][i] -= sh;
                    }
                }
            }
#pragma unroll
            for (int qb = 0; qb < 2; qb++) {
                // no key seen yet: weights exp2(-inf) = 0
                const S off = FOLD || m[qb] == -INFINITY ? S(0) : m[qb];
                S rk[4];
#pragma unroll
                for (int kb = 0; kb < 4; kb++) {
#pragma unroll
                    for (int i = 0; i < 4; i++) s[qb][kb][i] = fast_exp2(FOLD ? s[qb][kb][i] : s[qb][kb][i] - off);
                    rk[kb] = (s[qb][kb][0] + s[qb][kb][1]) + (s[qb][kb][2] + s[qb][kb][3]);
                }
                l[qb] += group_sum((rk[0] + rk[1]) + (rk[2] + rk[3]));
            }
            // O^T += V^T P^T, P^T's 32-key chunks from two stacked S blocks.
#pragma unroll
            for (int kc = 0; kc < 2; kc++) {
                Frag<P> pf[2];
#pragma unroll
                for (int qb = 0; qb < 2; qb++) {
                    S tmp[8];
#pragma unroll
                    for (int i = 0; i < 4; i++) { tmp[i] = s[qb][2 * kc][i]; tmp[i + 4] = s[qb][2 * kc + 1][i]; }
                    make_frag<P>(pf[qb], tmp);
                }
#pragma unroll
                for (int dh = 0; dh < DB; dh += 2) {
                    Frag<P> vf[2];
#pragma unroll
                    for (int db = 0; db < 2; db++)
                        tile_load_cperm<P>(vf[db], vt, F::VPLANE, ((dh + db) * 16 + c) * F::VP + kc * 32, g);
#pragma unroll
                    for (int sp_ = 0; sp_ < NP; sp_++)
#pragma unroll
                        for (int db = 0; db < 2; db++)
#pragma unroll
                            for (int qb = 0; qb < 2; qb++) mma_part<P>(o[qb][dh + db], vf[db], pf[qb], sp_);
                }
            }
            // This wave's loads have landed; after the barrier, everyone's
            // have, and no wave still reads this buffer.
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            cur ^= 1;
        }
        // The sums rescaled from the reference max to the true max.
#pragma unroll
        for (int qb = 0; qb < 2; qb++) {
            const S beta = m[qb] == -INFINITY ? S(1) : fast_exp2(m[qb] - mt[qb]);
            m[qb] = mt[qb];
            l[qb] *= beta;
#pragma unroll
            for (int db = 0; db < DB; db++)
#pragma unroll
                for (int i = 0; i < 4; i++) o[qb][db][i] *= beta;
        }
        // out[x][d] for d = db*16 + crow(g, i): the weighted sums over the
        // normalizer (a query with no keys has l = 0 and gets 0/0 = NaN).
        if (a.parts == 2) {
#pragma unroll
            for (int qb = 0; qb < 2; qb++) {
                if (x[qb] >= a.tq) continue;
                const int64_t r = (part * a.batches + b) * a.tq + x[qb];
                S *PO = static_cast<S *>(a.po) + r * D;
#pragma unroll
                for (int db = 0; db < DB; db++)
#pragma unroll
                    for (int i = 0; i < 4; i++) PO[db * 16 + crow<P>(g, i)] = o[qb][db][i];
                if (g == 0) {
                    static_cast<S *>(a.pm)[r] = m[qb];
                    static_cast<S *>(a.pl)[r] = l[qb];
                }
            }
            continue;
        }
#pragma unroll
        for (int qb = 0; qb < 2; qb++) {
            if (x[qb] >= a.tq) continue;
            In *O = static_cast<In *>(a.out) + (b * a.tq + x[qb]) * D;
#pragma unroll
            for (int db = 0; db < DB; db++)
#pragma unroll
                for (int i = 0; i < 4; i++) O[db * 16 + crow<P>(g, i)] = (In)(o[qb][db][i] / l[qb]);
            if (g == 0) {
                typedef typename Stat<P>::T St;
                static_cast<St *>(a.row_max)[b * a.tq + x[qb]] = (St)(m[qb] * (S)LN2);
                static_cast<St *>(a.row_sum)[b * a.tq + x[qb]] = (St)l[qb];
            }
        }
    }
    if (tid == 0 && a.visited && visited) atomicAdd(a.visited, visited);
}
#undef EF_WAVE_FLAGS
#undef EF_LOAD_TILE

// The two parts of each query's keys combined, part 0 then part 1: the
// sums rescaled to the larger max (log2 units), then normalized.
template <class S, class In>
__global__ __launch_bounds__(256) void attn_merge_kernel(const S *po, const S *pm, const S *pl, int64_t rows, int d,
                                                         In *out, S *row_max, S *row_sum) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * d) return;
    const int64_t r = i / d;
    const S m0 = pm[r], m1 = pm[rows + r];
    const S mx = m0 > m1 ? m0 : m1;
    const S a0 = m0 == -INFINITY ? S(0) : fast_exp2(m0 - mx);
    const S a1 = m1 == -INFINITY ? S(0) : fast_exp2(m1 - mx);
    const S l = pl[r] * a0 + pl[rows + r] * a1;
    out[i] = (In)((po[i] * a0 + po[rows * d + i] * a1) / l);
    if (i % d == 0) {
        row_max[r] = mx * (S)LN2;
        row_sum[r] = l;
    }
}

// out = (p0 + p1) * scale, part 0 first; a second array (p2, scale2, out2)
// of the same size in the same launch.
template <class S, class In>
__global__ __launch_bounds__(256) void attn_sum_kernel(const S *p, int64_t n, double scale, In *out, const S *p2,
                                                       double scale2, In *out2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {   // the second array, if any, in the grid's upper half
        i -= n;
        p = p2;
        scale = scale2;
        out = out2;
    }
    if (p && i < n) out[i] = (In)((p[i] + p[n + i]) * (S)scale);
}

// Any shape, one thread per (query, block of 8 value columns), keys in
// order: the reference's algorithm with a tile of one row.  Float64 or
// Float32 arithmetic throughout.
template <class T>
__global__ __launch_bounds__(256) void attn_fwd_generic(const T *Q, const T *K, const T *V, T *O,
                                                         T *row_max, T *row_sum, Mask mask,
                                                         int64_t batches, int64_t tq, int64_t tk,
                                                         int64_t d, int64_t dv, double scale) {
    constexpr int YC = 8;
    const int64_t chunks = (dv + YC - 1) / YC;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= batches * tq * chunks) return;
    const int64_t yc = t % chunks, x = (t / chunks) % tq, b = t / (chunks * tq);
    const Span sp = query_span(mask, x);
    const T *q = Q + (b * tq + x) * d;
    const T *kb = K + b * tk * d;
    const T *vb = V + b * tk * dv;
    T acc[YC];
    for (int i = 0; i < YC; i++) acc[i] = 0;
    T m = -INFINITY, z = 0;
    for (int64_t j = sp.lo; j < sp.hi; j++) {
        T s = 0;
        for (int64_t e = 0; e < d; e++) s += q[e] * kb[j * d + e];
        s *= (T)scale;
        if (s > m) {
            const T r = exp(m - s);
            z *= r;
            for (int i = 0; i < YC; i++) acc[i] *= r;
            m = s;
        }
        const T w = exp(s - m);
        z += w;
        for (int i = 0; i < YC; i++) {
            const int64_t y = yc * YC + i;
            if (y < dv) acc[i] += w * vb[j * dv + y];
        }
    }
    for (int i = 0; i < YC; i++) {
        const int64_t y = yc * YC + i;
        if (y < dv) O[(b * tq + x) * dv + y] = acc[i] / z;
    }
    if (yc == 0) {
        row_max[b * tq + x] = m;
        row_sum[b * tq + x] = z;
    }
}

// ---------------------------------------------------------------------
// Backward
// ---------------------------------------------------------------------
struct BwdArgs {
    const void *q, *k, *v;             // originals (register fragments)
    const void *dout;
    const void *qr, *qt, *dor, *dot;   // prepared query-side tiles
    const void *kr, *kt, *vr;          // prepared key-side tiles
    const void *row_max, *row_sum, *delta;   // Stat<P>::T per query
    void *dq, *dk, *dv;
    Counter *ctr;
    Mask mask;
    int64_t tq, tk, tqp, tkp, batches;
    double scale;
    int items, reverse, causal_only, parts;
    void *pdk, *pdv, *pdq;             // parts == 2: each part's unscaled sums
};


// dK and dV for 128 keys per workgroup (32 per wave), visiting the query
// tiles that may see them.
template <int P, int D>
__global__ __launch_bounds__(THREADS, 1) void attn_bwd_dkdv_kernel(BwdArgs a) {
    typedef typename Input<P>::T In;
    typedef typename Acc<P>::type AccV;
    typedef typename Acc<P>::scalar S;
    typedef typename Elem<P>::T E;
    constexpr int PL = Elem<P>::planes;
    constexpr int KC = D / 32, DB = D / 16;
    constexpr int RP = D + RPAD<E>, TP = BN + 8;
    constexpr int RPLANE = BN * RP, TPLANE = D * TP;
    __shared__ __attribute__((aligned(16))) E qs[PL * RPLANE];
    __shared__ __attribute__((aligned(16))) E qtr[PL * TPLANE];
    __shared__ __attribute__((aligned(16))) E dos[PL * RPLANE];
    __shared__ __attribute__((aligned(16))) E dotr[PL * TPLANE];
    typedef typename Stat<P>::T St;
    __shared__ St sm[BN], sd[BN];   // per query: m log2(e) + log2(l), and delta
    __shared__ int slo[BN], shi[BN];   // spans in 32 bits (the host keeps tq, tk < 2^31)
    __shared__ int slot;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, c = lane & 15;
    (void)lane;
    const int64_t nkt = (a.tk + BW - 1) / BW;
    const int64_t nqt = (a.tq + BN - 1) / BN;
    const int64_t rplane = a.batches * a.tqp * D, tplane = a.batches * D * a.tqp;
    const S scale2 = (S)a.scale * (S)LOG2E;

    ItemQueue queue;
    queue.start(a.ctr, &slot, a.items);
    for (;;) {
        const int item = queue.next();
        if (item >= a.items) break;
        const int64_t b = item % a.batches;
        const int64_t rest = item / a.batches;
        const int part = (int)(rest % a.parts);
        const int64_t kt = a.reverse ? nkt - 1 - rest / a.parts : rest / a.parts;
        const int64_t k0 = kt * BW;
        const int64_t kvalid = (k0 + BW < a.tk ? k0 + BW : a.tk);   // keys past tk are padding
        const In *K = static_cast<const In *>(a.k) + b * a.tk * D;
        const In *V = static_cast<const In *>(a.v) + b * a.tk * D;
        const E *QR = static_cast<const E *>(a.qr) + b * a.tqp * D;
        const E *QT = static_cast<const E *>(a.qt) + b * D * a.tqp;
        const E *DR = static_cast<const E *>(a.dor) + b * a.tqp * D;
        const E *DT = static_cast<const E *>(a.dot) + b * D * a.tqp;
        const St *RM = static_cast<const St *>(a.row_max) + b * a.tq;
        const St *RS = static_cast<const St *>(a.row_sum) + b * a.tq;
        const St *DL = static_cast<const St *>(a.delta) + b * a.tq;

        int kj[2];
        Frag<P> kf[2][KC], vf[2][KC];
#pragma unroll
        for (int kb = 0; kb < 2; kb++) {
            kj[kb] = (int)k0 + w * 32 + kb * 16 + c;
#pragma unroll
            for (int kc = 0; kc < KC; kc++) {
                In t1[8], t2[8];
#pragma unroll
                for (int j = 0; j < 8; j++) {   // unconditional loads (row clamped), then 0 past tk
                    const int64_t kr = kj[kb] < a.tk ? kj[kb] : a.tk - 1;
                    t1[j] = K[kr * D + kc * 32 + g * 8 + j];
                    t2[j] = V[kr * D + kc * 32 + g * 8 + j];
                }
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    t1[j] = kj[kb] < a.tk ? t1[j] : In(0);
                    t2[j] = kj[kb] < a.tk ? t2[j] : In(0);
                }
                make_frag<P>(kf[kb][kc], t1);
                make_frag<P>(vf[kb][kc], t2);
            }
        }
        AccV dk[2][DB], dv[2][DB];
#pragma unroll
        for (int kb = 0; kb < 2; kb++)
#pragma unroll
            for (int i = 0; i < DB; i++) { dk[kb][i] = AccV{0, 0, 0, 0}; dv[kb][i] = AccV{0, 0, 0, 0}; }

        // Under a causal mask alone, query x sees key j when j <= x + offset.
        int64_t qt_begin = 0, qt_end = nqt;
        if (a.causal_only) {
            const int64_t xmin = k0 - a.mask.offset;
            qt_begin = xmin > 0 ? xmin / BN : 0;
            qt_begin = qt_begin < nqt ? qt_begin : nqt;
        }
        if (a.parts == 2) {
            const int64_t cut = qt_begin + (qt_end - qt_begin + 1) / 2;
            if (part == 0) qt_end = cut;
            else qt_begin = cut;
        }
        // The query tiles, the next one's operands and row statistics
        // loaded into registers while this one is computed.
        typedef TileIO<E, PL, BN, D> RowT;   // [query][feature]
        typedef TileIO<E, PL, D, BN> ColT;   // [feature][query]
        constexpr bool PF = 2 * (RowT::N + ColT::N) <= 16;
        u32x4 pq[RowT::N], pqt[ColT::N], pd[RowT::N], pdt[ColT::N];
        St prm = 0, prs = 1, pdl = 0;
#define EF_FETCH(T_)                                                                         \
    do {                                                                                     \
        const int64_t q0_ = (T_) * BN;                                                       \
        /* row statistics first: loads complete in order, so using them */                  \
        /* later waits for nothing behind them */                                            \
        if (tid < BN) {                                                                      \
            const int64_t xc_ = q0_ + tid < a.tq ? q0_ + tid : a.tq - 1;                     \
            prm = RM[xc_];                                                                   \
            prs = RS[xc_];                                                                   \
            pdl = DL[xc_];                                                                   \
        }                                                                                    \
        RowT::load(pq, QR + q0_ * D, D, rplane);                                             \
        ColT::load(pqt, QT + q0_, a.tqp, tplane);                                            \
        RowT::load(pd, DR + q0_ * D, D, rplane);                                             \
        ColT::load(pdt, DT + q0_, a.tqp, tplane);                                            \
    } while (0)
        // Part k of the same loads for tile T, so that the pipelined tile
        // spreads them over its steps: 0 the row statistics, then the chunks
        // of Q rows, Q^T, dO rows and dO^T in turn.
        constexpr int FETCH_PARTS = 1 + 2 * (RowT::N + ColT::N);
        auto fetch_part = [&](int64_t T, int k) __attribute__((always_inline)) {
            const int64_t q0_ = T * BN;
            if (k == 0) {
                if (tid < BN) {
                    const int64_t xc_ = q0_ + tid < a.tq ? q0_ + tid : a.tq - 1;
                    prm = RM[xc_];
                    prs = RS[xc_];
                    pdl = DL[xc_];
                }
                return;
            }
            int i = k - 1;
            if (i < RowT::N) { RowT::load_chunk(pq, i, QR + q0_ * D, D, rplane); return; }
            i -= RowT::N;
            if (i < ColT::N) { ColT::load_chunk(pqt, i, QT + q0_, a.tqp, tplane); return; }
            i -= ColT::N;
            if (i < RowT::N) { RowT::load_chunk(pd, i, DR + q0_ * D, D, rplane); return; }
            i -= RowT::N;
            ColT::load_chunk(pdt, i, DT + q0_, a.tqp, tplane);
        };
        // FP64: each half's stages one after another, the mask tested
        // unless the tile is full.
        auto dkdv_tile_f64 = [&](bool full) __attribute__((always_inline)) {
#pragma unroll 1
                for (int hc = 0; hc < 2; hc++) {
                    AccV s1[2][2], dp1[2][2];   // [query block][key block]
#pragma unroll
                    for (int qq = 0; qq < 2; qq++)
#pragma unroll
                        for (int kb = 0; kb < 2; kb++) { s1[qq][kb] = AccV{0, 0, 0, 0}; dp1[qq][kb] = AccV{0, 0, 0, 0}; }
#pragma unroll
                    for (int kc = 0; kc < KC; kc++)
#pragma unroll
                        for (int qq = 0; qq < 2; qq++) {
                            const int qb = 2 * hc + qq;
                            Frag<P> af, df;
                            tile_load8<P>(af, qs, RPLANE, (qb * 16 + c) * RP + kc * 32 + g * 8);
                            tile_load8<P>(df, dos, RPLANE, (qb * 16 + c) * RP + kc * 32 + g * 8);
#pragma unroll
                            for (int sp_ = 0; sp_ < Parts<P>::n; sp_++)
#pragma unroll
                                for (int kb = 0; kb < 2; kb++) {
                                    mma_part<P>(s1[qq][kb], af, kf[kb][kc], sp_);
                                    mma_part<P>(dp1[qq][kb], df, vf[kb][kc], sp_);
                                }
                        }
                    S p[2][2][4], ds[2][2][4];
#pragma unroll
                    for (int qq = 0; qq < 2; qq++)
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const int r = (2 * hc + qq) * 16 + crow<P>(g, i);
                            const S mr = (S)sm[r], dr = (S)sd[r];
                            const int lo = slo[r], hi = shi[r];
#pragma unroll
                            for (int kb = 0; kb < 2; kb++) {
                                S x = s1[qq][kb][i] * scale2 - mr;
                                if (!full) x = ((kj[kb] >= lo) & (kj[kb] < hi)) ? x : -INFINITY;
                                const S pv = fast_exp2(x);
                                p[qq][kb][i] = pv;
                                ds[qq][kb][i] = pv * (dp1[qq][kb][i] - dr);
                            }
                        }
                    Frag<P> pf1[2], dsf1[2];
#pragma unroll
                    for (int kb = 0; kb < 2; kb++) {
                        S t1[8], t2[8];
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            t1[i] = p[0][kb][i]; t1[i + 4] = p[1][kb][i];
                            t2[i] = ds[0][kb][i]; t2[i + 4] = ds[1][kb][i];
                        }
                        make_frag<P>(pf1[kb], t1);
                        make_frag<P>(dsf1[kb], t2);
                    }
#pragma unroll
                    for (int db = 0; db < DB; db++) {
                        Frag<P> bo, bq;
                        const int row = (db * 16 + c) * TP + hc * 32;
                        tile_load_cperm<P>(bo, dotr, TPLANE, row, g);
                        tile_load_cperm<P>(bq, qtr, TPLANE, row, g);
#pragma unroll
                        for (int sp_ = 0; sp_ < Parts<P>::n; sp_++)
#pragma unroll
                            for (int kb = 0; kb < 2; kb++) {
                                mma_part<P>(dv[kb][db], pf1[kb], bo, sp_);
                                mma_part<P>(dk[kb][db], dsf1[kb], bq, sp_);
                            }
                    }
                }
        };
        // One query tile's share of dK and dV, over its two 32-query halves
        // h: S = Q K^T and dP = dO V^T (MFMA), then P and dS (VALU), then
        // dV += P^T dO and dK += dS^T Q (MFMA).  FULL: every query of the
        // tile sees every key of the block (no mask test).
        //
        // Below FP64 the stages run as a pipeline inside each wave, in
        // steps that each pair a group of MFMAs with vector work that does
        // not depend on them: S and dP of half 1 beside P and dS of half 0,
        // dV and dK of half 0 beside P and dS of half 1.  An MFMA runs for
        // many cycles after its issue, and the wave issues the step's VALU
        // instructions meanwhile.  Each step's LDS operands are read one
        // step ahead; a scheduling barrier closes each step, so the
        // compiler keeps this order.
        auto dkdv_tile = [&](auto full_c, int64_t next) __attribute__((always_inline)) {
            constexpr bool FULL = decltype(full_c)::value;
            AccV s[2][2][2], dp[2][2][2];   // [half][query block][key block]
            Frag<P> pf[2][2], dsf[2][2];    // [half][key block]
            // Row statistics of half hc, query block qq: rows crow(g, 0..3)
            // are consecutive, so one 16-byte read per array (float rows).
            St mrow[2][2][4], drow[2][2][4];
            int lorow[2][2][4], hirow[2][2][4];
            auto load_stats = [&](int hc) __attribute__((always_inline)) {
#pragma unroll
                for (int qq = 0; qq < 2; qq++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int r = (2 * hc + qq) * 16 + crow<P>(g, i);
                        mrow[hc][qq][i] = sm[r];
                        drow[hc][qq][i] = sd[r];
                        if constexpr (!FULL) { lorow[hc][qq][i] = slo[r]; hirow[hc][qq][i] = shi[r]; }
                    }
            };
            auto load_ad = [&](int hc, int kc, int qq, Frag<P> &af, Frag<P> &df) __attribute__((always_inline)) {
                const int qb = 2 * hc + qq;
                tile_load8<P>(af, qs, RPLANE, (qb * 16 + c) * RP + kc * 32 + g * 8);
                tile_load8<P>(df, dos, RPLANE, (qb * 16 + c) * RP + kc * 32 + g * 8);
            };
            auto mma_sd = [&](int hc, int kc, int qq, const Frag<P> &af, const Frag<P> &df) __attribute__((always_inline)) {
                if (kc == 0) {
#pragma unroll
                    for (int kb = 0; kb < 2; kb++) { s[hc][qq][kb] = AccV{0, 0, 0, 0}; dp[hc][qq][kb] = AccV{0, 0, 0, 0}; }
                }
#pragma unroll
                for (int sp_ = 0; sp_ < Parts<P>::n; sp_++)
#pragma unroll
                    for (int kb = 0; kb < 2; kb++) {
                        mma_part<P>(s[hc][qq][kb], af, kf[kb][kc], sp_);
                        mma_part<P>(dp[hc][qq][kb], df, vf[kb][kc], sp_);
                    }
            };
            auto load_bo = [&](int hc, int db, Frag<P> &bo, Frag<P> &bq) __attribute__((always_inline)) {
                const int row = (db * 16 + c) * TP + hc * 32;
                tile_load_cperm<P>(bo, dotr, TPLANE, row, g);
                tile_load_cperm<P>(bq, qtr, TPLANE, row, g);
            };
            auto mma_kv = [&](int hc, int db, const Frag<P> &bo, const Frag<P> &bq) __attribute__((always_inline)) {
#pragma unroll
                for (int sp_ = 0; sp_ < Parts<P>::n; sp_++)
#pragma unroll
                    for (int kb = 0; kb < 2; kb++) {
                        mma_part<P>(dv[kb][db], pf[hc][kb], bo, sp_);
                        mma_part<P>(dk[kb][db], dsf[hc][kb], bq, sp_);
                    }
            };
            // P = exp2(scale s log2(e) - m log2(e) - log2(l)) and dS =
            // P (dP - delta) for rows crow(g, 2 ip), crow(g, 2 ip + 1) of
            // query block qq, into fragment slots 4 qq + 2 ip, + 1.  A
            // masked score gets exponent -inf: weight 0, no branch.
            auto soft = [&](int hc, int qq, int ip) __attribute__((always_inline)) {
#pragma unroll
                for (int kb = 0; kb < 2; kb++) {
                    S pv[2], dv_[2];
#pragma unroll
                    for (int t = 0; t < 2; t++) {
                        const int i = 2 * ip + t;
                        S x = s[hc][qq][kb][i] * scale2 - (S)mrow[hc][qq][i];
                        if constexpr (!FULL)
                            x = ((kj[kb] >= lorow[hc][qq][i]) & (kj[kb] < hirow[hc][qq][i])) ? x : -INFINITY;
                        pv[t] = fast_exp2(x);
                        dv_[t] = pv[t] * (dp[hc][qq][kb][i] - (S)drow[hc][qq][i]);
                    }
                    set_pair<P>(pf[hc][kb], 4 * qq + 2 * ip, pv[0], pv[1]);
                    set_pair<P>(dsf[hc][kb], 4 * qq + 2 * ip, dv_[0], dv_[1]);
                }
            };
            {
                // Steps: S, dP of half 0 (phase A) and of half 1 (B), then
                // dV, dK of half 0 (C) and of half 1 (D).  Step n + 1's LDS
                // operands are read during step n.
                constexpr int NA = 2 * KC, NC = DB, N = 2 * NA + 2 * NC;
                auto load_step = [&](int n, Frag<P> &u, Frag<P> &v) __attribute__((always_inline)) {
                    if (n < 2 * NA) load_ad(n / NA, (n % NA) / 2, n % 2, u, v);
                    else load_bo((n - 2 * NA) / NC, (n - 2 * NA) % NC, u, v);
                };
                auto mma_step = [&](int n, const Frag<P> &u, const Frag<P> &v) __attribute__((always_inline)) {
                    if (n < 2 * NA) mma_sd(n / NA, (n % NA) / 2, n % 2, u, v);
                    else mma_kv((n - 2 * NA) / NC, (n - 2 * NA) % NC, u, v);
                };
                // P and dS of a half in four pieces (query block, row pair),
                // spread evenly over the NS steps of the phase beside it.
                auto pieces = [&](int hc, int m, int NS) __attribute__((always_inline)) {
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (k * NS / 4 == m) soft(hc, k / 2, k % 2);
                };
                Frag<P> u[2], v[2];
                load_stats(0);
                load_step(0, u[0], v[0]);
#pragma unroll
                for (int n = 0; n < N; n++) {
                    if (n + 1 < N) load_step(n + 1, u[(n + 1) & 1], v[(n + 1) & 1]);
                    mma_step(n, u[n & 1], v[n & 1]);
                    if (n >= NA && n < 2 * NA) pieces(0, n - NA, NA);
                    if (n >= 2 * NA && n < 2 * NA + NC) pieces(1, n - 2 * NA, NC);
                    if (n == NA - 1) load_stats(1);
                    // The next tile's loads (with PF), a few per step of
                    // the last phase (pure MFMA, the fewest live values):
                    // the memory pipeline takes them while the wave computes
                    // instead of holding the wave at one burst of 17.
                    if constexpr (PF) {
#pragma unroll
                        for (int k = 0; k < FETCH_PARTS; k++)
                            if (2 * NA + NC + k * NC / FETCH_PARTS == n) fetch_part(next, k);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        };
        // Tiles [qa, qb): FULL (a compile-time choice) runs the pipeline
        // without mask tests and needs every query of each tile to see every
        // key of the block; otherwise each tile's flags choose.
        auto run_tiles = [&](auto full_c, int64_t qa, int64_t qb) __attribute__((always_inline)) {
            constexpr bool FULL = decltype(full_c)::value;
            for (int64_t qt = qa; qt < qb; qt++) {
                const int64_t q0 = qt * BN;
                if (!PF) EF_FETCH(qt);
                RowT::store(qs, RP, RPLANE, pq);
                ColT::store(qtr, TP, TPLANE, pqt);
                RowT::store(dos, RP, RPLANE, pd);
                ColT::store(dotr, TP, TPLANE, pdt);
                // Every wave votes over the tile's 64 queries (one per lane),
                // so all agree without a barrier: does any see a key of the
                // block, does each see all of them.
                bool any_tile, full_tile;
                {
                    const int64_t xq = q0 + lane;
                    const Span sp = query_span(a.mask, xq);
                    any_tile = FULL || __any(sp.lo < sp.hi && sp.lo < kvalid && sp.hi > k0);
                    full_tile = FULL || __all(xq >= a.tq || (sp.lo <= k0 && sp.hi >= kvalid));
                    if (tid < BN) {
                        slo[tid] = (int)sp.lo;
                        shi[tid] = (int)sp.hi;
                        sm[tid] = xq < a.tq ? prm * (St)LOG2E + log2_of(prs) : St(0);
                        sd[tid] = xq < a.tq ? pdl : St(0);
                    }
                }
                __syncthreads();
                // The pipelined tile issues the next tile's loads itself (the
                // last tile reloads its own: harmless, and no branch).
                const int64_t nxt = qt + 1 < qt_end ? qt + 1 : qt;
                if constexpr (P == PREC_F64) {
                    if (PF && qt + 1 < qt_end) EF_FETCH(qt + 1);
                    if (any_tile) dkdv_tile_f64(full_tile);
                } else if constexpr (FULL) {
                    dkdv_tile(std::true_type{}, nxt);
                } else {
                    if (!any_tile) {
                        if (PF && qt + 1 < qt_end) EF_FETCH(qt + 1);
                    } else if (full_tile) {
                        dkdv_tile(std::true_type{}, nxt);
                    } else {
                        dkdv_tile(std::false_type{}, nxt);
                    }
                }
                __syncthreads();
            }
        };
        // Where every tile from some point on is full (a causal mask alone:
        // from the first tile whose first query sees the block's last key;
        // no mask: all), those tiles run in a loop of their own, so the
        // loop's code holds one variant of the pipeline.
        int64_t qt_full = qt_end;
        if (!a.mask.allowed && !a.mask.causal) {
            qt_full = qt_begin;
        } else if (a.causal_only) {
            const int64_t need = kvalid - 1 - a.mask.offset;   // first query that sees key kvalid - 1
            const int64_t t = need <= 0 ? 0 : (need + BN - 1) / BN;
            qt_full = t < qt_begin ? qt_begin : (t > qt_end ? qt_end : t);
        }
        if (PF && qt_begin < qt_end) EF_FETCH(qt_begin);
        if constexpr (P == PREC_F64) {
            run_tiles(std::false_type{}, qt_begin, qt_end);
        } else {
            run_tiles(std::false_type{}, qt_begin, qt_full);
            run_tiles(std::true_type{}, qt_full, qt_end);
        }
#undef EF_FETCH
        // Lane holds dK/dV[key k0 + w*32 + kb*16 + crow(g, i)][d = db*16 + c].
        if (a.parts == 2) {
            S *PK = static_cast<S *>(a.pdk) + (part * a.batches + b) * a.tk * D;
            S *PV = static_cast<S *>(a.pdv) + (part * a.batches + b) * a.tk * D;
#pragma unroll
            for (int kb = 0; kb < 2; kb++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int64_t key = k0 + w * 32 + kb * 16 + crow<P>(g, i);
                    if (key < a.tk) {
#pragma unroll
                        for (int db = 0; db < DB; db++) {
                            PK[key * D + db * 16 + c] = dk[kb][db][i];
                            PV[key * D + db * 16 + c] = dv[kb][db][i];
                        }
                    }
                }
            continue;
        }
        In *DK = static_cast<In *>(a.dk) + b * a.tk * D;
        In *DV = static_cast<In *>(a.dv) + b * a.tk * D;
#pragma unroll
        for (int kb = 0; kb < 2; kb++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t key = k0 + w * 32 + kb * 16 + crow<P>(g, i);
                if (key < a.tk) {
#pragma unroll
                    for (int db = 0; db < DB; db++) {
                        DK[key * D + db * 16 + c] = (In)(dk[kb][db][i] * (S)a.scale);
                        DV[key * D + db * 16 + c] = (In)dv[kb][db][i];
                    }
                }
            }
    }
}

// dQ for 128 queries per workgroup (32 per wave), visiting the key tiles
// each may see.
template <int P, int D>
__global__ __launch_bounds__(THREADS, 1) void attn_bwd_dq_kernel(BwdArgs a) {
    typedef typename Input<P>::T In;
    typedef typename Acc<P>::type AccV;
    typedef typename Acc<P>::scalar S;
    typedef typename Elem<P>::T E;
    constexpr int PL = Elem<P>::planes;
    constexpr int KC = D / 32, DB = D / 16;
    constexpr int RP = D + RPAD<E>, TP = BN + 8;
    constexpr int RPLANE = BN * RP, TPLANE = D * TP;
    __shared__ __attribute__((aligned(16))) E ks[PL * RPLANE];
    __shared__ __attribute__((aligned(16))) E ktr[PL * TPLANE];
    __shared__ __attribute__((aligned(16))) E vs[PL * RPLANE];
    __shared__ int slot;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, c = lane & 15;
    (void)lane;
    const int64_t nqt = (a.tq + BW - 1) / BW;
    const int64_t rplane = a.batches * a.tkp * D, tplane = a.batches * D * a.tkp;
    const S scale2 = (S)a.scale * (S)LOG2E;

    ItemQueue queue;
    queue.start(a.ctr, &slot, a.items);
    for (;;) {
        const int item = queue.next();
        if (item >= a.items) break;
        const int64_t b = item % a.batches;
        const int64_t rest = item / a.batches;
        const int part = (int)(rest % a.parts);
        const int64_t qt = a.reverse ? nqt - 1 - rest / a.parts : rest / a.parts;
        const int64_t q0 = qt * BW;
        const In *Q = static_cast<const In *>(a.q) + b * a.tq * D;
        const In *DO = static_cast<const In *>(a.dout) + b * a.tq * D;
        const E *KR = static_cast<const E *>(a.kr) + b * a.tkp * D;
        const E *KT = static_cast<const E *>(a.kt) + b * D * a.tkp;
        const E *VR = static_cast<const E *>(a.vr) + b * a.tkp * D;

        int64_t x[2];
        Span sp[2];
        int slo32[2], shi32[2];   // spans in 32 bits (tk < 2^31)
        Frag<P> qf[2][KC], df[2][KC];
        S mrow[2], drow[2];
#pragma unroll
        for (int qb = 0; qb < 2; qb++) {
            x[qb] = q0 + w * 32 + qb * 16 + c;
            sp[qb] = query_span(a.mask, x[qb]);
            slo32[qb] = (int)sp[qb].lo;
            shi32[qb] = (int)sp[qb].hi;
#pragma unroll
            for (int kc = 0; kc < KC; kc++) {
                In t1[8], t2[8];
#pragma unroll
                for (int j = 0; j < 8; j++) {   // unconditional loads (row clamped), then 0 past tq
                    const int64_t xr = x[qb] < a.tq ? x[qb] : a.tq - 1;
                    t1[j] = Q[xr * D + kc * 32 + g * 8 + j];
                    t2[j] = DO[xr * D + kc * 32 + g * 8 + j];
                }
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    t1[j] = x[qb] < a.tq ? t1[j] : In(0);
                    t2[j] = x[qb] < a.tq ? t2[j] : In(0);
                }
                make_frag<P>(qf[qb][kc], t1);
                make_frag<P>(df[qb][kc], t2);
            }
            // m log2(e) + log2(l): P = exp2(scale s log2(e) - mrow).
            {
                typedef typename Stat<P>::T St;
                const int64_t xr = b * a.tq + (x[qb] < a.tq ? x[qb] : a.tq - 1);
                const St rm = static_cast<const St *>(a.row_max)[xr], rsum = static_cast<const St *>(a.row_sum)[xr];
                const St dl = static_cast<const St *>(a.delta)[xr];
                mrow[qb] = x[qb] < a.tq ? (S)(rm * (St)LOG2E + log2_of(rsum)) : S(0);
                drow[qb] = x[qb] < a.tq ? (S)dl : S(0);
            }
        }
        int64_t first, last;
        block_range(a.mask, q0, first, last);
        int64_t jbeg = first < last ? first - first % BN : last;
        if (a.parts == 2) {
            const int64_t cut = jbeg + ((last - jbeg + 2 * BN - 1) / (2 * BN)) * BN;
            if (part == 0) last = cut < last ? cut : last;
            else jbeg = cut < last ? cut : last;
        }
        AccV dq[2][DB];
#pragma unroll
        for (int qb = 0; qb < 2; qb++)
#pragma unroll
            for (int i = 0; i < DB; i++) dq[qb][i] = AccV{0, 0, 0, 0};

        // The key tiles, the next one's loaded into registers while this one
        // is computed.
        typedef TileIO<E, PL, BN, D> RowT;   // [key][feature]
        typedef TileIO<E, PL, D, BN> ColT;   // [feature][key]
        constexpr bool PF = 2 * RowT::N + ColT::N <= 16;
        u32x4 pk[RowT::N], pkt[ColT::N], pv_[RowT::N];
#define EF_FETCH(J_)                                                                         \
    do {                                                                                     \
        RowT::load(pk, KR + (J_) * D, D, rplane);                                            \
        ColT::load(pkt, KT + (J_), a.tkp, tplane);                                           \
        RowT::load(pv_, VR + (J_) * D, D, rplane);                                           \
    } while (0)
        // Chunk k of the same loads for the tile at J, spread by the
        // pipelined tile over its last phase: K rows, K^T, V rows in turn.
        constexpr int FETCH_PARTS = 2 * RowT::N + ColT::N;
        auto fetch_part = [&](int64_t J, int k) __attribute__((always_inline)) {
            if (k < RowT::N) { RowT::load_chunk(pk, k, KR + J * D, D, rplane); return; }
            k -= RowT::N;
            if (k < ColT::N) { ColT::load_chunk(pkt, k, KT + J, a.tkp, tplane); return; }
            k -= ColT::N;
            RowT::load_chunk(pv_, k, VR + J * D, D, rplane);
        };
        // FP64: each half's stages one after another, the mask tested
        // unless the tile is full.
        auto dq_tile_f64 = [&](bool full_tile, int64_t j0) __attribute__((always_inline)) {
            // The tile's keys in two 32-key halves: S^T = K Q^T and
            // dP^T = V dO^T, then dS^T, then dQ^T += K^T dS^T.
#pragma unroll 1
            for (int hc = 0; hc < 2; hc++) {
                AccV s[2][2], dp[2][2];   // [key block][query block]
#pragma unroll
                for (int kk = 0; kk < 2; kk++)
#pragma unroll
                    for (int qb = 0; qb < 2; qb++) { s[kk][qb] = AccV{0, 0, 0, 0}; dp[kk][qb] = AccV{0, 0, 0, 0}; }
#pragma unroll
                for (int kc = 0; kc < KC; kc++) {
#pragma unroll
                    for (int kk = 0; kk < 2; kk++) {
                        const int kb = 2 * hc + kk;
                        Frag<P> af, vf;
                        tile_load8<P>(af, ks, RPLANE, (kb * 16 + c) * RP + kc * 32 + g * 8);
                        tile_load8<P>(vf, vs, RPLANE, (kb * 16 + c) * RP + kc * 32 + g * 8);
#pragma unroll
                        for (int sp_ = 0; sp_ < Parts<P>::n; sp_++)
#pragma unroll
                            for (int qb = 0; qb < 2; qb++) {
                                mma_part<P>(s[kk][qb], af, qf[qb][kc], sp_);
                                mma_part<P>(dp[kk][qb], vf, df[qb][kc], sp_);
                            }
                    }
                }
                Frag<P> dsf[2];
#pragma unroll
                for (int qb = 0; qb < 2; qb++) {
                    S t[8];
                    if (full_tile) {
#pragma unroll
                        for (int kk = 0; kk < 2; kk++)
#pragma unroll
                            for (int i = 0; i < 4; i++) {
                                const S pv = fast_exp2(s[kk][qb][i] * scale2 - mrow[qb]);
                                t[kk * 4 + i] = pv * (dp[kk][qb][i] - drow[qb]);
                            }
                    } else {
#pragma unroll
                        for (int kk = 0; kk < 2; kk++)
#pragma unroll
                            for (int i = 0; i < 4; i++) {
                                const int64_t j = j0 + (2 * hc + kk) * 16 + crow<P>(g, i);
                                const S e = fast_exp2(s[kk][qb][i] * scale2 - mrow[qb]);
                                const S pv = j >= sp[qb].lo && j < sp[qb].hi ? e : S(0);
                                t[kk * 4 + i] = pv * (dp[kk][qb][i] - drow[qb]);
                            }
                    }
                    make_frag<P>(dsf[qb], t);
                }
                Frag<P> kt[DB];
#pragma unroll
                for (int db = 0; db < DB; db++) tile_load_cperm<P>(kt[db], ktr, TPLANE, (db * 16 + c) * TP + hc * 32, g);
#pragma unroll
                for (int sp_ = 0; sp_ < Parts<P>::n; sp_++)
#pragma unroll
                    for (int db = 0; db < DB; db++)
#pragma unroll
                        for (int qb = 0; qb < 2; qb++) mma_part<P>(dq[qb][db], kt[db], dsf[qb], sp_);
            }
        };
        // One key tile's share of dQ, over its two 32-key halves h: S^T =
        // K Q^T and dP^T = V dO^T (MFMA), then dS^T (VALU), then dQ^T +=
        // K^T dS^T (MFMA), as a pipeline of steps inside each wave (the
        // dK/dV kernel's schedule): S^T and dP^T of half 1 beside dS^T of
        // half 0, dQ^T of half 0 beside dS^T of half 1.  FULL: every query
        // of the wave sees every key of the tile.
        auto dq_tile = [&](auto full_c, int64_t j0, int64_t next) __attribute__((always_inline)) {
            constexpr bool FULL = decltype(full_c)::value;
            AccV s[2][2][2], dp[2][2][2];   // [half][key block][query block]
            Frag<P> dsf[2][2];              // [half][query block]
            auto load_kv = [&](int hc, int kc, int kk, Frag<P> &af, Frag<P> &vf) __attribute__((always_inline)) {
                const int kb = 2 * hc + kk;
                tile_load8<P>(af, ks, RPLANE, (kb * 16 + c) * RP + kc * 32 + g * 8);
                tile_load8<P>(vf, vs, RPLANE, (kb * 16 + c) * RP + kc * 32 + g * 8);
            };
            auto mma_sd = [&](int hc, int kc, int kk, const Frag<P> &af, const Frag<P> &vf) __attribute__((always_inline)) {
                if (kc == 0) {
#pragma unroll
                    for (int qb = 0; qb < 2; qb++) { s[hc][kk][qb] = AccV{0, 0, 0, 0}; dp[hc][kk][qb] = AccV{0, 0, 0, 0}; }
                }
#pragma unroll
                for (int sp_ = 0; sp_ < Parts<P>::n; sp_++)
#pragma unroll
                    for (int qb = 0; qb < 2; qb++) {
                        mma_part<P>(s[hc][kk][qb], af, qf[qb][kc], sp_);
                        mma_part<P>(dp[hc][kk][qb], vf, df[qb][kc], sp_);
                    }
            };
            auto load_kt = [&](int hc, int db, Frag<P> &kt) __attribute__((always_inline)) {
                tile_load_cperm<P>(kt, ktr, TPLANE, (db * 16 + c) * TP + hc * 32, g);
            };
            auto mma_dq = [&](int hc, int db, const Frag<P> &kt) __attribute__((always_inline)) {
#pragma unroll
                for (int sp_ = 0; sp_ < Parts<P>::n; sp_++)
#pragma unroll
                    for (int qb = 0; qb < 2; qb++) mma_part<P>(dq[qb][db], kt, dsf[hc][qb], sp_);
            };
            // dS^T of key block kk, query block qb of half hc: fragment
            // slots 4 kk .. 4 kk + 3.  A masked score gets exponent -inf.
            auto soft = [&](int hc, int qb, int kk) __attribute__((always_inline)) {
                S t[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    S xs = s[hc][kk][qb][i] * scale2 - mrow[qb];
                    if constexpr (!FULL) {
                        const int j = (int)j0 + (2 * hc + kk) * 16 + crow<P>(g, i);
                        xs = ((j >= slo32[qb]) & (j < shi32[qb])) ? xs : -INFINITY;
                    }
                    t[i] = fast_exp2(xs) * (dp[hc][kk][qb][i] - drow[qb]);
                }
                set_pair<P>(dsf[hc][qb], 4 * kk, t[0], t[1]);
                set_pair<P>(dsf[hc][qb], 4 * kk + 2, t[2], t[3]);
            };
            constexpr int NA = 2 * KC, NC = DB, N = 2 * NA + 2 * NC;
            auto load_step = [&](int n, Frag<P> &u, Frag<P> &v) __attribute__((always_inline)) {
                if (n < 2 * NA) load_kv(n / NA, (n % NA) / 2, n % 2, u, v);
                else load_kt((n - 2 * NA) / NC, (n - 2 * NA) % NC, u);
            };
            auto mma_step = [&](int n, const Frag<P> &u, const Frag<P> &v) __attribute__((always_inline)) {
                if (n < 2 * NA) mma_sd(n / NA, (n % NA) / 2, n % 2, u, v);
                else mma_dq((n - 2 * NA) / NC, (n - 2 * NA) % NC, u);
            };
            auto pieces = [&](int hc, int m, int NS) __attribute__((always_inline)) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (k * NS / 4 == m) soft(hc, k / 2, k % 2);
            };
            Frag<P> u[2], v[2];
            load_step(0, u[0], v[0]);
#pragma unroll
            for (int n = 0; n < N; n++) {
                if (n + 1 < N) load_step(n + 1, u[(n + 1) & 1], v[(n + 1) & 1]);
                mma_step(n, u[n & 1], v[n & 1]);
                if (n >= NA && n < 2 * NA) pieces(0, n - NA, NA);
                if (n >= 2 * NA && n < 2 * NA + NC) pieces(1, n - 2 * NA, NC);
                // The next tile's loads (with PF), a few per step of the last
                // phase, instead of one burst after the barrier.
                if constexpr (PF) {
#pragma unroll
                    for (int k = 0; k < FETCH_PARTS; k++)
                        if (2 * NA + NC + k * NC / FETCH_PARTS == n) fetch_part(next, k);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        // Key tiles [ja, jb): FULL (a compile-time choice) needs every query
        // of the workgroup to see every key of each tile.
        auto run_tiles = [&](auto full_c, int64_t ja, int64_t jb) __attribute__((always_inline)) {
            constexpr bool FULL = decltype(full_c)::value;
            for (int64_t j0 = ja; j0 < jb; j0 += BN) {
                if (!PF) EF_FETCH(j0);
                RowT::store(ks, RP, RPLANE, pk);
                ColT::store(ktr, TP, TPLANE, pkt);
                RowT::store(vs, RP, RPLANE, pv_);
                // Whether every query of the wave sees every real key of the
                // tile (then no mask test per element).
                bool full_tile = true;
                if constexpr (!FULL) {
                    const int64_t jvalid = j0 + BN < a.tk ? j0 + BN : a.tk;
                    bool full = true;
#pragma unroll
                    for (int qb = 0; qb < 2; qb++) full &= x[qb] >= a.tq || (sp[qb].lo <= j0 && sp[qb].hi >= jvalid);
                    full_tile = __all(full);
                }
                __syncthreads();
                // The pipelined tile issues the next tile's loads itself (the
                // last tile reloads its own: harmless, and no branch).
                const int64_t nxt = j0 + BN < last ? j0 + BN : j0;
                if constexpr (P == PREC_F64) {
                    if (PF && j0 + BN < last) EF_FETCH(j0 + BN);
                    dq_tile_f64(full_tile, j0);
                } else if constexpr (FULL) {
                    dq_tile(std::true_type{}, j0, nxt);
                } else if (full_tile) {
                    dq_tile(std::true_type{}, j0, nxt);
                } else {
                    dq_tile(std::false_type{}, j0, nxt);
                }
                __syncthreads();
            }
        };
        // Under a causal mask alone the first key tiles are seen whole by
        // every query of the block (a tile is, when its last key is at most
        // q0 + offset); with no mask, all are.  Those run in a loop of their
        // own that holds one variant of the pipeline.
        int64_t j_full = jbeg;
        if (!a.mask.allowed && !a.mask.causal) {
            j_full = last;
        } else if (a.causal_only) {
            const int64_t lim = q0 + a.mask.offset + 1;   // query q0 sees keys below lim
            if (lim >= a.tk) {
                j_full = last;
            } else {
                const int64_t n = lim - BN - jbeg;
                j_full = n < 0 ? jbeg : jbeg + (n / BN + 1) * BN;
            }
            j_full = j_full < jbeg ? jbeg : (j_full > last ? last : j_full);
        }
        if (PF && jbeg < last) EF_FETCH(jbeg);
        if constexpr (P == PREC_F64) {
            run_tiles(std::false_type{}, jbeg, last);
        } else {
            run_tiles(std::true_type{}, jbeg, j_full);
            run_tiles(std::false_type{}, j_full, last);
        }
#undef EF_FETCH
        // Lane holds dQ^T[d = db*16 + crow(g, i)][query x].
        if (a.parts == 2) {
#pragma unroll
            for (int qb = 0; qb < 2; qb++) {
                if (x[qb] >= a.tq) continue;
                S *PQ = static_cast<S *>(a.pdq) + ((part * a.batches + b) * a.tq + x[qb]) * D;
#pragma unroll
                for (int db = 0; db < DB; db++)
#pragma unroll
                    for (int i = 0; i < 4; i++) PQ[db * 16 + crow<P>(g, i)] = dq[qb][db][i];
            }
            continue;
        }
#pragma unroll
        for (int qb = 0; qb < 2; qb++) {
            if (x[qb] >= a.tq) continue;
            In *DQ = static_cast<In *>(a.dq) + (b * a.tq + x[qb]) * D;
#pragma unroll
            for (int db = 0; db < DB; db++)
#pragma unroll
                for (int i = 0; i < 4; i++) DQ[db * 16 + crow<P>(g, i)] = (In)(dq[qb][db][i] * (S)a.scale);
        }
    }
}

Mask make_mask(const ef_attention_desc &d) {
    Mask m;
    m.allowed = d.allowed;
    m.causal = d.causal;
    m.offset = d.causal_offset;
    m.window = d.window;
    m.tq = (int64_t)d.queries;
    m.tk = (int64_t)d.keys;
    return m;
}

int compute_units() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = n;
        else
            cus = 256;
    }
    return cus;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

bool fast_fwd_shape(const ef_attention_desc &d) {
    if (d.features != d.values) return false;
    // The kernel's 32-bit offsets into a plane of prepared operands.
    if (2 * d.batches * (size_t)round_up((int64_t)d.keys, PAD) * d.features >= (size_t(1) << 31)) return false;
    if (d.queries >= (size_t(1) << 31) - BW) return false;   // 32-bit query rows
    if (d.precision == EF_PRECISION_FP64) return d.features == 32 || d.features == 64;
    return d.features == 32 || d.features == 64 || d.features == 128;
}

// Two parts per block where the mask makes blocks' work differ and there
// are too few blocks to even it out over the CUs.
int parts_for(const ef_attention_desc &d, int64_t blocks) {
    const bool varying = d.allowed || d.causal;
    return varying && blocks < 4 * (int64_t)compute_units() ? 2 : 1;
}


bool fast_bwd_shape(const ef_attention_desc &d) {
    // The tile loads' 32-bit byte offsets: up to 32 rows of padded queries
    // (8-byte elements) from a common base.
    if ((size_t)round_up((int64_t)d.queries, PAD) >= (size_t(1) << 24)) return false;
    return d.features == d.values && (d.features == 32 || d.features == 64);
}

size_t elem_bytes(uint32_t precision) {
    switch (precision) {
    case EF_PRECISION_FP64: return 8;
    case EF_PRECISION_FP32: return 4;
    case EF_PRECISION_BF16X3: return 4;   // two BF16 planes
    default: return 2;
    }
}

// Scratch layout, each region 256-byte aligned: the work counters, then
// the forward's prepared K and V^T and its two parts' partial results
// (sums, max and sum), or (from the same offset) the backward's prepared
// tiles, each query's delta and its parts' dK, dV and dQ.  The forward
// needs only its own regions (fwd_total bytes).
struct Layout {
    size_t counters, kr, vt;
    size_t qr, qt, dor, dot, bkr, bkt, bvr, delta;
    size_t po, pm, pl, pdk, pdv, pdq;
    size_t fwd_total, total;
};

Layout layout(const ef_attention_desc &d) {
    Layout L;
    const size_t e = elem_bytes(d.precision);
    const size_t tqp = (size_t)round_up((int64_t)d.queries, PAD), tkp = (size_t)round_up((int64_t)d.keys, PAD);
    const size_t qa = align256(d.batches * tqp * d.features * e), ka = align256(d.batches * tkp * d.features * e);
    const size_t sb = d.precision == EF_PRECISION_FP64 ? 8 : 4;
    const size_t rq = 2 * d.batches * d.queries, rk = 2 * d.batches * d.keys;
    size_t at = 0;
    L.counters = at; at += 256;
    L.kr = at; at += ka;
    L.vt = at; at += ka;
    L.po = at; at += align256(rq * d.values * sb);
    L.pm = at; at += align256(rq * sb);
    L.pl = at; at += align256(rq * sb);
    L.fwd_total = at;
    at = 256;
    L.qr = at; at += qa;
    L.qt = at; at += qa;
    L.dor = at; at += qa;
    L.dot = at; at += qa;
    L.bkr = at; at += ka;
    L.bkt = at; at += ka;
    L.bvr = at; at += ka;
    L.delta = at; at += align256(d.batches * d.queries * sb);
    L.pdk = at; at += align256(rk * d.features * sb);
    L.pdv = at; at += align256(rk * d.values * sb);
    L.pdq = at; at += align256(rq * d.features * sb);
    L.total = at > L.fwd_total ? at : L.fwd_total;
    return L;
}

template <int P, int D>
ef_status run_fwd(const ef_attention_desc &d, const void *q, const void *k, const void *v, void *out,
                  void *row_max, void *row_sum, uint64_t *visited, char *scratch, hipStream_t s) {
    typedef typename Input<P>::T In;
    typedef typename Elem<P>::T E;
    const Layout L = layout(d);
    const int64_t B = (int64_t)d.batches, tq = (int64_t)d.queries, tk = (int64_t)d.keys;
    const int64_t tkp = round_up(tk, PAD);
    E *kr = reinterpret_cast<E *>(scratch + L.kr);
    E *vt = reinterpret_cast<E *>(scratch + L.vt);
    PrepArgs pa = {};
    pa.job[0] = {k, tk, tkp, kr, nullptr};
    pa.job[1] = {v, tk, tkp, nullptr, vt};
    pa.njobs = 2;
    pa.B = B;
    pa.zero = reinterpret_cast<Counter *>(scratch + L.counters);
    pa.nzero = QUEUES;
    hipLaunchKernelGGL((prep_kernel<P, D>), dim3((unsigned)(tkp / 64), (unsigned)B, 2), dim3(256), 0, s, pa);
    FwdArgs a;
    a.q = q; a.kr = kr; a.vt = vt; a.out = out;
    a.row_max = row_max; a.row_sum = row_sum;
    a.visited = reinterpret_cast<unsigned long long *>(visited);
    a.mask = make_mask(d);
    a.tq = tq; a.tk = tk; a.tkp = tkp; a.batches = B;
    a.scale = d.scale;
    a.ctr = reinterpret_cast<Counter *>(scratch + L.counters);
    const int64_t blocks = B * ((tq + BW - 1) / BW);
    a.parts = parts_for(d, blocks);
    a.pad = 0;
    a.items = (int)(blocks * a.parts);
    a.reverse = d.causal ? 1 : 0;
    a.po = scratch + L.po; a.pm = scratch + L.pm; a.pl = scratch + L.pl;
    const int cap = 2 * compute_units();
    const int grid = a.items < cap ? a.items : cap;
    hipLaunchKernelGGL((attn_fwd_kernel<P, D>), dim3((unsigned)grid), dim3(THREADS), 0, s, a);
    if (a.parts == 2) {
        typedef typename Acc<P>::scalar S;
        const int64_t rows = B * tq;
        hipLaunchKernelGGL((attn_merge_kernel<S, In>), dim3(blocks_for(rows * D, 256)), dim3(256), 0, s,
                           (const S *)a.po, (const S *)a.pm, (const S *)a.pl, rows, D, (In *)out, (S *)row_max,
                           (S *)row_sum);
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

template <int P>
ef_status dispatch_fwd(const ef_attention_desc &d, const void *q, const void *k, const void *v, void *out,
                       void *m, void *l, uint64_t *vis, char *scratch, hipStream_t s) {
    switch (d.features) {
    case 32: return run_fwd<P, 32>(d, q, k, v, out, m, l, vis, scratch, s);
    case 64: return run_fwd<P, 64>(d, q, k, v, out, m, l, vis, scratch, s);
    case 128:
        if constexpr (P != PREC_F64) return run_fwd<P, 128>(d, q, k, v, out, m, l, vis, scratch, s);
        [[fallthrough]];
    default:
        set_error("no fast attention kernel for this head size");
        return EF_INVALID;
    }
}

template <int P, int D>
ef_status run_bwd(const ef_attention_desc &d, const void *q, const void *k, const void *v, const void *out,
                  const void *dout, const void *row_max, const void *row_sum, void *dq, void *dk, void *dv,
                  char *scratch, int stages, hipStream_t s) {
    typedef typename Input<P>::T In;
    typedef typename Elem<P>::T E;
    const Layout L = layout(d);
    const int64_t B = (int64_t)d.batches, tq = (int64_t)d.queries, tk = (int64_t)d.keys;
    const int64_t tqp = round_up(tq, PAD), tkp = round_up(tk, PAD);
    BwdArgs a;
    a.q = q; a.k = k; a.v = v; a.dout = dout;
    a.qr = scratch + L.qr; a.qt = scratch + L.qt; a.dor = scratch + L.dor; a.dot = scratch + L.dot;
    a.kr = scratch + L.bkr; a.kt = scratch + L.bkt; a.vr = scratch + L.bvr;
    a.row_max = row_max; a.row_sum = row_sum;
    a.delta = scratch + L.delta;
    a.dq = dq; a.dk = dk; a.dv = dv;
    a.mask = make_mask(d);
    a.tq = tq; a.tk = tk; a.tqp = tqp; a.tkp = tkp; a.batches = B;
    a.scale = d.scale;
    a.causal_only = (!d.allowed && d.causal && d.window <= 0) ? 1 : 0;
    a.pdk = scratch + L.pdk; a.pdv = scratch + L.pdv; a.pdq = scratch + L.pdq;
    typedef typename Acc<P>::scalar S;
    if (stages & 1) {
        // Q, dO, K and V prepared, delta computed and both work counters
        // zeroed, in one launch.
        PrepArgs pa = {};
        pa.job[0] = {q, tq, tqp, scratch + L.qr, scratch + L.qt};
        pa.job[1] = {dout, tq, tqp, scratch + L.dor, scratch + L.dot};
        pa.job[2] = {k, tk, tkp, scratch + L.bkr, scratch + L.bkt};
        pa.job[3] = {v, tk, tkp, scratch + L.bvr, nullptr};
        pa.njobs = 4;
        pa.B = B;
        pa.zero = reinterpret_cast<Counter *>(scratch + L.counters);
        pa.nzero = 2 * QUEUES;
        pa.out = out;
        pa.dout = dout;
        pa.delta = scratch + L.delta;
        pa.tq = tq;
        const int64_t xb = (tqp > tkp ? tqp : tkp) / 64;
        hipLaunchKernelGGL((prep_kernel<P, D>), dim3((unsigned)xb, (unsigned)B, 5), dim3(256), 0, s, pa);
    }
    const int cap = 2 * compute_units();
    if (stages & 2) {
        a.ctr = reinterpret_cast<Counter *>(scratch + L.counters);
        const int64_t blocks = B * ((tk + BW - 1) / BW);
        a.parts = parts_for(d, blocks);
        a.items = (int)(blocks * a.parts);
        a.reverse = 0;   // key tiles in order: under a causal mask the first are the heaviest
        const int grid = a.items < cap ? a.items : cap;
        if (!(stages & 1)) hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, a.ctr);
        attn_bwd_launch(0, P, D, &a, (unsigned)grid, s);
        if (a.parts == 2) {
            const int64_t n = B * tk * D;
            hipLaunchKernelGGL((attn_sum_kernel<S, In>), dim3(blocks_for(2 * n, 256)), dim3(256), 0, s,
                               (const S *)a.pdk, n, d.scale, (In *)dk, (const S *)a.pdv, 1.0, (In *)dv);
        }
    }
    if (stages & 4) {
        a.ctr = reinterpret_cast<Counter *>(scratch + L.counters) + QUEUES;
        const int64_t blocks = B * ((tq + BW - 1) / BW);
        a.parts = parts_for(d, blocks);
        a.items = (int)(blocks * a.parts);
        a.reverse = d.causal ? 1 : 0;   // under a causal mask the last query tiles are the heaviest
        const int grid = a.items < cap ? a.items : cap;
        if (!(stages & 1)) hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, a.ctr);
        attn_bwd_launch(1, P, D, &a, (unsigned)grid, s);
        if (a.parts == 2) {
            const int64_t n = B * tq * D;
            hipLaunchKernelGGL((attn_sum_kernel<S, In>), dim3(blocks_for(n, 256)), dim3(256), 0, s, (const S *)a.pdq, n,
                               d.scale, (In *)dq, (const S *)nullptr, 0.0, (In *)nullptr);
        }
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

template <int P>
ef_status dispatch_bwd(const ef_attention_desc &d, const void *q, const void *k, const void *v, const void *out,
                       const void *dout, const void *m, const void *l, void *dq, void *dk, void *dv, char *scratch,
                       int stages, hipStream_t s) {
    switch (d.features) {
    case 32: return run_bwd<P, 32>(d, q, k, v, out, dout, m, l, dq, dk, dv, scratch, stages, s);
    case 64: return run_bwd<P, 64>(d, q, k, v, out, dout, m, l, dq, dk, dv, scratch, stages, s);
    default:
        set_error("the attention backward pass takes 32 or 64 features, equal for keys and values");
        return EF_INVALID;
    }
}

}  // namespace

#if EF_ATTN_BWD_TU
template <int P, int D> void bwd_launch_one(int kernel, const void *args, unsigned grid, hipStream_t s) {
    const BwdArgs a = *static_cast<const BwdArgs *>(args);
    if (kernel == 0) hipLaunchKernelGGL((attn_bwd_dkdv_kernel<P, D>), dim3(grid), dim3(THREADS), 0, s, a);
    else hipLaunchKernelGGL((attn_bwd_dq_kernel<P, D>), dim3(grid), dim3(THREADS), 0, s, a);
}

template <int P> void bwd_launch_p(int kernel, int d, const void *args, unsigned grid, hipStream_t s) {
    if (d == 32) bwd_launch_one<P, 32>(kernel, args, grid, s);
    else bwd_launch_one<P, 64>(kernel, args, grid, s);
}

void attn_bwd_launch(int kernel, int precision, int d, const void *args, unsigned grid, hipStream_t s) {
    switch (precision) {
    case PREC_F64: bwd_launch_p<PREC_F64>(kernel, d, args, grid, s); break;
    case PREC_F32: bwd_launch_p<PREC_F32>(kernel, d, args, grid, s); break;
    case PREC_BF16X3: bwd_launch_p<PREC_BF16X3>(kernel, d, args, grid, s); break;
    default: bwd_launch_p<PREC_BF16>(kernel, d, args, grid, s); break;
    }
}
#else

size_t attention_scratch_bytes(const ef_attention_desc &d) { return layout(d).total; }

size_t attention_fwd_scratch_bytes(const ef_attention_desc &d) {
    // The generic kernel (shapes the MFMA kernels don't take) uses none.
    if (d.batches == 0 || d.queries == 0 || !(fast_fwd_shape(d) && d.precision <= EF_PRECISION_BF16)) return 0;
    return layout(d).fwd_total;
}

ef_status attention_fwd(const ef_attention_desc &d, const void *q, const void *k, const void *v, void *out,
                        void *row_max, void *row_sum, uint64_t *tiles_visited, void *scratch, hipStream_t s) {
    if (d.batches == 0 || d.queries == 0) return EF_OK;
    if (fast_fwd_shape(d) && d.precision <= EF_PRECISION_BF16) {
        if (!scratch) {
            set_error("the attention forward pass needs its scratch");
            return EF_INVALID;
        }
        char *sc = static_cast<char *>(scratch);
        switch (d.precision) {
        case EF_PRECISION_FP32: return dispatch_fwd<PREC_F32>(d, q, k, v, out, row_max, row_sum, tiles_visited, sc, s);
        case EF_PRECISION_BF16X3: return dispatch_fwd<PREC_BF16X3>(d, q, k, v, out, row_max, row_sum, tiles_visited, sc, s);
        case EF_PRECISION_BF16: return dispatch_fwd<PREC_BF16>(d, q, k, v, out, row_max, row_sum, tiles_visited, sc, s);
        default: return dispatch_fwd<PREC_F64>(d, q, k, v, out, row_max, row_sum, tiles_visited, sc, s);
        }
    }
    // Any other shape: the generic kernel (FP64 for double tensors, FP32
    // for float ones; no tile count).
    const Mask mask = make_mask(d);
    const int64_t chunks = ((int64_t)d.values + 7) / 8;
    const size_t threads = d.batches * d.queries * (size_t)chunks;
    if (d.precision == EF_PRECISION_FP64)
        hipLaunchKernelGGL(attn_fwd_generic<double>, dim3(blocks_for(threads, 256)), dim3(256), 0, s,
                           (const double *)q, (const double *)k, (const double *)v, (double *)out, (double *)row_max,
                           (double *)row_sum,
                           mask, (int64_t)d.batches, (int64_t)d.queries, (int64_t)d.keys, (int64_t)d.features,
                           (int64_t)d.values, d.scale);
    else
        hipLaunchKernelGGL(attn_fwd_generic<float>, dim3(blocks_for(threads, 256)), dim3(256), 0, s,
                           (const float *)q, (const float *)k, (const float *)v, (float *)out, (float *)row_max,
                           (float *)row_sum,
                           mask, (int64_t)d.batches, (int64_t)d.queries, (int64_t)d.keys, (int64_t)d.features,
                           (int64_t)d.values, d.scale);
    EF_HIP(hipGetLastError());
    return EF_OK;
}

ef_status attention_bwd(const ef_attention_desc &d, const void *q, const void *k, const void *v, const void *out,
                        const void *dout, const void *row_max, const void *row_sum, void *dq, void *dk, void *dv,
                        void *scratch, int stages, hipStream_t s) {
    if (d.batches == 0) return EF_OK;
    if (!fast_bwd_shape(d)) {
        set_error("the attention backward pass takes 32 or 64 features, equal for keys and values");
        return EF_INVALID;
    }
    if (!scratch) {
        set_error("the attention backward pass needs its scratch");
        return EF_INVALID;
    }
    char *sc = static_cast<char *>(scratch);
    switch (d.precision) {
    case EF_PRECISION_FP32: return dispatch_bwd<PREC_F32>(d, q, k, v, out, dout, row_max, row_sum, dq, dk, dv, sc, stages, s);
    case EF_PRECISION_BF16X3: return dispatch_bwd<PREC_BF16X3>(d, q, k, v, out, dout, row_max, row_sum, dq, dk, dv, sc, stages, s);
    case EF_PRECISION_BF16: return dispatch_bwd<PREC_BF16>(d, q, k, v, out, dout, row_max, row_sum, dq, dk, dv, sc, stages, s);
    case EF_PRECISION_FP64: return dispatch_bwd<PREC_F64>(d, q, k, v, out, dout, row_max, row_sum, dq, dk, dv, sc, stages, s);
    default:
        set_error("unknown precision");
        return EF_INVALID;
    }
}

#endif  // EF_ATTN_BWD_TU

}  // namespace ef
