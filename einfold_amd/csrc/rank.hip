// Ranking and sorting on gfx950 (crates/einfold-kernels/src/
// discriminate.rs:103-156, 280-292): each row's rank among the distinct
// words and each rank's first row, and a stable sort of (word, row) pairs.
//
// Ranking from 65536 rows first deduplicates the words through a hash
// table and orders only the distinct ones (the hashing path, below the
// radix kernels); where that cannot finish (more distinct words than a
// quarter of the rows, or words crowded into a narrow part of their range)
// it ranks by the radix sort.  Both give the same ranks and first rows:
// integer results that depend on no order of execution.
//
// The radix sort: a stable least-significant-byte-first sort of
// (word, row) pairs that skips the bytes no word varies in, then each
// row's rank among the distinct words and each rank's first row.
//
// One read of the words builds all eight byte histograms; a small launch
// adds up the workgroups' histograms, from which the host learns which
// bytes vary.  Each varying byte is then one pass of three launches
// (reduce, then scan, then scatter): each tile of 4096 pairs counts its
// values of the byte; a scan over the tiles of each value gives every tile
// the start of its run of each value; and each tile
//   1. ranks its pairs among the tile's pairs with the same byte, in row
//      order: a wave takes 1024 consecutive rows, 64 at a time, finds the
//      lanes that share its byte with eight ballots and counts them in a
//      histogram of its own in LDS (no barrier inside the wave);
//   2. stages its pairs in LDS in sorted order and writes each value's run
//      to its place, so that neighbouring lanes write neighbouring words.
// (A single-pass scatter with a decoupled look-back across tiles, as in
// Adinets and Merrill's "Onesweep", 2022, measured slower here: every tile
// waits for the slowest tile before it.)
// The ranks come the same way: each tile of sorted positions counts the
// positions that start a rank (word differs from the one before) and,
// for large inputs, its rows in each of up to 256 buckets of rows; a scan
// gives each tile its first rank and its runs of (row, rank) pairs; one
// workgroup per bucket then gathers its pairs in an LDS window and writes
// the bucket's ranks in whole lines (4-byte stores to random rows would
// cost HBM a 32-byte sector each).  All of it is integer arithmetic: exact
// and deterministic.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.h"

namespace ef {

namespace {

constexpr int THREADS = 256;
constexpr int WAVES = THREADS / 64;
constexpr int KPT = 16;                     // pairs per thread per tile
constexpr int TILE = THREADS * KPT;         // 4096 pairs per tile
constexpr int WAVE_PAIRS = 64 * KPT;        // consecutive rows of one wave

constexpr int HIST_CHUNKS = 64;             // workgroups per byte adding the histograms up

// Small device words of a call (the number of ranks is word S_RANKS).
constexpr int SMALL_WORDS = 64;
constexpr int S_RANKS = 10;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

struct HistArgs {
    const uint64_t *w;
    int64_t n, tiles;
    uint32_t *partial;          // [tiles][8][256]
    unsigned *small;            // SMALL_WORDS words to clear
    uint32_t *counts;           // [8][256], cleared here, summed by the next launch
};

// The eight byte histograms of tile blockIdx.x, in LDS.  A byte's
// histograms are also the tile counts of the first pass, which reads the
// words in this order.
__global__ __launch_bounds__(THREADS) void byte_histograms_kernel(HistArgs a) {
    __shared__ uint32_t h[8][256];
    const int tid = threadIdx.x, lane = tid & 63;
#pragma unroll
    for (int b = 0; b < 8; b++) h[b][tid] = 0;
    if (blockIdx.x == 0) {
        if (tid < SMALL_WORDS) a.small[tid] = 0;
#pragma unroll
        for (int b = 0; b < 8; b++) a.counts[b * 256 + tid] = 0;
    }
    const int64_t base = (int64_t)blockIdx.x * TILE;
    uint64_t v[KPT];
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int64_t i = base + r * THREADS + tid;
        v[r] = a.w[i < a.n ? i : a.n - 1];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const bool valid = base + r * THREADS + tid < a.n;
        const uint64_t live = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint32_t d = (uint32_t)(v[r] >> (8 * b)) & 0xff;
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
            // One add for the wave when every live lane has the same byte
            // (live lanes are a prefix of the wave, so lane 0 is one of
            // them).
            if (__ballot(valid && d == d0) == live) {
                if (lane == 0 && live) atomicAdd(&h[b][d0], (uint32_t)__popcll(live));
            } else if (valid) {
                atomicAdd(&h[b][d], 1u);
            }
        }
    }
    __syncthreads();
    uint32_t *out = a.partial + (int64_t)blockIdx.x * 2048;
#pragma unroll
    for (int b = 0; b < 8; b++) out[b * 256 + tid] = h[b][tid];
}

// counts[b][d] = words whose byte b is d: workgroup (c, b) adds up
// histogram workgroups c, c + HIST_CHUNKS, ... (integer sums: the order
// does not matter).
__global__ __launch_bounds__(THREADS) void sum_histograms_kernel(const uint32_t *partial, int blocks,
                                                                 uint32_t *counts) {
    const int tid = threadIdx.x, c = blockIdx.x, b = blockIdx.y;
    const uint32_t *col = partial + b * 256 + tid;
    uint32_t s = 0;
    int k = c;
    for (; k + 7 * HIST_CHUNKS < blocks; k += 8 * HIST_CHUNKS) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = col[(int64_t)(k + j * HIST_CHUNKS) * 2048];
#pragma unroll
        for (int j = 0; j < 8; j++) s += v[j];
    }
    for (; k < blocks; k += HIST_CHUNKS) s += col[(int64_t)k * 2048];
    if (s) atomicAdd(&counts[b * 256 + tid], s);
}

// Tile t's count of each value of the byte at `shift`: tile_counts[d][t].
template <class W>
__global__ __launch_bounds__(THREADS) void tile_counts_kernel(const W *__restrict__ w, int64_t n, int shift,
                                                              uint32_t *__restrict__ tile_counts, int64_t tiles) {
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x, lane = tid & 63;
    h[tid] = 0;
    const int64_t base = (int64_t)blockIdx.x * TILE;
    W v[KPT];
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int64_t i = base + r * THREADS + tid;
        v[r] = w[i < n ? i : n - 1];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const bool valid = base + r * THREADS + tid < n;
        const uint64_t live = __ballot(valid);
        const uint32_t d = (uint32_t)(v[r] >> shift) & 0xff;
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        // One add for the wave when every live lane has the same value
        // (live lanes are a prefix of the wave).
        if (__ballot(valid && d == d0) == live) {
            if (lane == 0 && live) atomicAdd(&h[d0], (uint32_t)__popcll(live));
        } else if (valid) {
            atomicAdd(&h[d], 1u);
        }
    }
    __syncthreads();
    tile_counts[(int64_t)tid * tiles + blockIdx.x] = h[tid];
}

// Output start of tile t's run of value d: the words of values below d,
// plus the words of value d in tiles t' < t (tile_counts[d * sd + t' * st]).
// The words of each value are
// counts[value], or, with counts null, those of bucket `value` of the rows
// 0 .. n-1 in buckets of `bucket` rows.  Workgroup d, thread j takes tiles
// [j*per, (j+1)*per).
__global__ __launch_bounds__(THREADS) void tile_offsets_kernel(const uint32_t *__restrict__ tile_counts,
                                                               const uint32_t *__restrict__ counts, int64_t tiles,
                                                               uint32_t *__restrict__ offsets, int64_t n,
                                                               int64_t bucket, int64_t sd, int64_t st) {
    __shared__ uint32_t wt[2][WAVES];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, d = blockIdx.x;
    // words of values below d
    uint32_t of_tid;
    if (counts) {
        of_tid = counts[tid];
    } else {
        const int64_t left = n - tid * bucket;
        of_tid = (uint32_t)(left <= 0 ? 0 : (left < bucket ? left : bucket));
    }
    const uint32_t below_d = wave_sum(tid < d ? of_tid : 0u);
    const int64_t per = (tiles + THREADS - 1) / THREADS;
    const uint32_t *row = tile_counts + (int64_t)d * sd;
    const int64_t t0 = tid * per, t1 = t0 + per < tiles ? t0 + per : tiles;
    uint32_t mine = 0;
    for (int64_t t = t0; t < t1; t++) mine += row[t * st];
    const uint32_t incl = wave_inclusive_sum(mine, lane);
    if (lane == 63) {
        wt[0][wv] = incl;
        wt[1][wv] = below_d;
    }
    __syncthreads();
    uint32_t run = incl - mine, base = 0;
    for (int w = 0; w < WAVES; w++) {
        if (w < wv) run += wt[0][w];
        base += wt[1][w];
    }
    run += base;
    uint32_t *out = offsets + (int64_t)d * tiles;
    for (int64_t t = t0; t < t1; t++) {
        out[t] = run;
        run += row[t * st];
    }
}

struct SweepArgs {
    const uint64_t *w_in;
    const uint32_t *r_in;       // null: rows 0 .. n-1
    uint64_t *w_out;
    uint32_t *r_out;
    int64_t n;
    int shift;
    const uint32_t *offsets;    // [256][tiles] output start of each tile's run of each value
    int64_t tiles;
};

// One stable pass over the byte at `shift`, tile blockIdx.x.
__global__ __launch_bounds__(THREADS) void sweep_kernel(SweepArgs a) {
    // 52224 bytes of LDS: three workgroups per CU.  The scan's wave totals
    // live in `dst` until `dst` is written (after another barrier).
    __shared__ uint64_t sk[TILE];
    __shared__ uint32_t sr[TILE];
    __shared__ uint16_t cnt[WAVES][256];   // per wave: its pairs of each value, then their place (< TILE)
    __shared__ uint32_t dst[256];          // output position of the tile's value-d run, less its LDS start
    uint32_t *const wt = dst;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t tile = blockIdx.x;
    const uint32_t out_d = a.offsets[(int64_t)tid * a.tiles + tile];   // where this tile's run of value tid starts
#pragma unroll
    for (int k = 0; k < 4; k++) cnt[wv][k * 64 + lane] = 0;
    const int64_t base = tile * TILE + wv * WAVE_PAIRS;

    uint64_t key[KPT];
    uint32_t row[KPT];
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int64_t i = base + r * 64 + lane;
        key[r] = a.w_in[i < a.n ? i : a.n - 1];
    }
    if (a.r_in) {
#pragma unroll
        for (int r = 0; r < KPT; r++) {
            const int64_t i = base + r * 64 + lane;
            row[r] = a.r_in[i < a.n ? i : a.n - 1];
        }
    } else {
#pragma unroll
        for (int r = 0; r < KPT; r++) row[r] = (uint32_t)(base + r * 64 + lane);
    }

    // 1. Place of each pair among the wave's earlier pairs of its value.
    const uint64_t below = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    uint32_t off[KPT];
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const bool valid = base + r * 64 + lane < a.n;
        const uint32_t d = (uint32_t)(key[r] >> a.shift) & 0xff;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const bool on = (d >> bit) & 1;
            const uint64_t set = __ballot(on);
            peers &= on ? set : ~set;
        }
        const uint32_t pre = cnt[wv][d];
        // The highest lane of each group of peers advances the count; the
        // wave's LDS operations complete in order, so the next round reads
        // the new count.
        if (valid && (peers >> lane) == 1ULL) cnt[wv][d] = (uint16_t)(pre + (uint32_t)__popcll(peers));
        __builtin_amdgcn_wave_barrier();
        off[r] = pre + (uint32_t)__popcll(peers & below);
    }
    __syncthreads();

    // 2. Thread d: the tile's LDS start of each wave's run of value d.
    const int d = tid;
    uint32_t c[WAVES], total = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        c[w] = cnt[w][d];
        total += c[w];
    }
    const uint32_t incl = wave_inclusive_sum(total, lane);
    if (lane == 63) wt[wv] = incl;
    __syncthreads();
    uint32_t start = incl - total;
    for (int w = 0; w < wv; w++) start += wt[w];
    const uint32_t tile_start = start;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        cnt[w][d] = (uint16_t)start;
        start += c[w];
    }
    __syncthreads();

    // 3. The pairs to LDS in sorted order.
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const bool valid = base + r * 64 + lane < a.n;
        const uint32_t lp = cnt[wv][(uint32_t)(key[r] >> a.shift) & 0xff] + off[r];
        if (valid) {
            sk[lp] = key[r];
            sr[lp] = row[r];
        }
    }
    dst[d] = out_d - tile_start;
    __syncthreads();

    // 4. Runs out: LDS slot i goes to dst[value] + i.
    const int64_t t0 = tile * TILE;
    const int len = (int)(a.n - t0 < TILE ? a.n - t0 : TILE);
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int i = r * THREADS + tid;
        const uint64_t k = sk[i];
        const uint32_t rw = sr[i];
        const uint32_t pos = dst[(uint32_t)(k >> a.shift) & 0xff] + (uint32_t)i;
        if (i < len) {
            a.w_out[pos] = k;
            a.r_out[pos] = rw;
        }
    }
}

// Per tile of sorted positions: how many start a rank (a position whose
// word differs from the one before), to tile_counts[256][t]; with BUCKETS
// also its rows in each bucket of 2^bshift rows, to tile_counts[b][t].
template <bool BUCKETS>
__global__ __launch_bounds__(THREADS) void starts_kernel(const uint64_t *__restrict__ w,
                                                         const uint32_t *__restrict__ rows, int64_t n, int bshift,
                                                         uint32_t *__restrict__ tile_counts, int64_t tiles) {
    __shared__ uint32_t h[BUCKETS ? 256 : 1];
    __shared__ uint32_t heads_s;
    const int tid = threadIdx.x, lane = tid & 63;
    if (BUCKETS) h[tid] = 0;
    if (tid == 0) heads_s = 0;
    const int64_t t0 = (int64_t)blockIdx.x * TILE + tid, last = n - 1;
    uint64_t w0[KPT], w1[KPT];
    uint32_t rw[KPT];
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int64_t p = t0 + r * THREADS;
        const int64_t pc = p < last ? p : last;
        w0[r] = w[pc];
        w1[r] = w[pc > 0 ? pc - 1 : 0];
        if (BUCKETS) rw[r] = rows[pc];
    }
    __syncthreads();
    uint32_t starts = 0;
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int64_t p = t0 + r * THREADS;
        starts += p <= last && (p == 0 || w0[r] != w1[r]);
        if (BUCKETS && p <= last) atomicAdd(&h[rw[r] >> bshift], 1u);
    }
    starts = wave_sum(starts);
    if (lane == 0) atomicAdd(&heads_s, starts);
    __syncthreads();
    if (BUCKETS) tile_counts[(int64_t)tid * tiles + blockIdx.x] = h[tid];
    if (tid == 0) tile_counts[(int64_t)256 * tiles + blockIdx.x] = heads_s;
}

struct AssignArgs {
    const uint64_t *w;          // sorted words
    const uint32_t *rows;       // their rows
    int64_t n;
    int64_t tiles;
    uint32_t *rank, *first;
    const uint32_t *offsets;    // [257][tiles]: rows 0-255 bucket runs, row 256 ranks started before each tile
    uint32_t *n_ranks;
    int bshift;
    uint64_t *pairs;            // with BUCKETS: (row << 32 | rank) by buckets of 2^bshift rows
};

// Ranks from the sorted pairs: a position starts a rank where its word
// differs from the one before; the rank of a position is the number of
// starts up to it, less one.  A tile's positions are taken 256 at a time
// (round r, wave, lane), so loads are coalesced; the starts are counted
// per wave and round with ballots.  Without BUCKETS each rank goes
// straight to its row.  With BUCKETS the (row, rank) pairs are staged in
// LDS by bucket of rows (in any order inside a bucket) and written as runs,
// for rows_from_buckets_kernel: 4-byte stores to random rows cost HBM a
// 32-byte sector each.
constexpr int PARTS = KPT * WAVES;
static_assert(PARTS <= 64, "one wave scans the start counts of a tile's (round, wave) groups");

template <bool BUCKETS>
__global__ __launch_bounds__(THREADS) void assign_kernel(AssignArgs a) {
    __shared__ uint32_t part[PARTS];   // starts per (round, wave), then their exclusive prefix
    __shared__ uint64_t sp[BUCKETS ? TILE : 1];   // staged pairs
    __shared__ uint32_t bc[BUCKETS ? 256 : 1];    // pairs of each bucket, then the bucket's LDS start
    __shared__ uint32_t bdst[BUCKETS ? 256 : 1];  // output position of a bucket's run, less its LDS start
    __shared__ uint32_t bwt[WAVES];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t tile = blockIdx.x;
    if (BUCKETS) bc[tid] = 0;
    const uint32_t before = a.offsets[(int64_t)256 * a.tiles + tile];   // ranks started before this tile
    const uint32_t out_b = BUCKETS ? a.offsets[(int64_t)tid * a.tiles + tile] : 0;
    const int64_t t0 = tile * TILE + tid;
    const int64_t last = a.n - 1;
    const uint64_t below = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    uint64_t heads[KPT];
    uint32_t row[KPT];
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int64_t p = t0 + r * THREADS;
        const int64_t pc = p < last ? p : last;
        const uint64_t w = a.w[pc];
        const uint64_t prev = a.w[pc > 0 ? pc - 1 : 0];
        row[r] = a.rows[pc];
        heads[r] = __ballot(p <= last && (p == 0 || w != prev));
    }
    __syncthreads();
    uint32_t loc[BUCKETS ? KPT : 1];
    if (BUCKETS) {
#pragma unroll
        for (int r = 0; r < KPT; r++)
            loc[r] = t0 + r * THREADS <= last ? atomicAdd(&bc[row[r] >> a.bshift], 1u) : 0u;
    }
    if (lane < KPT) {
        uint64_t h = heads[0];
#pragma unroll
        for (int r = 1; r < KPT; r++) h = lane == r ? heads[r] : h;
        part[lane * WAVES + wv] = (uint32_t)__popcll(h);
    }
    __syncthreads();
    uint32_t bcount = 0, bincl = 0;
    if (BUCKETS) {
        bcount = bc[tid];
        bincl = wave_inclusive_sum(bcount, lane);
        if (lane == 63) bwt[wv] = bincl;
    }
    if (wv == 0) {
        // One wave scans the KPT * WAVES counts; lanes past them add 0 and
        // leave the neighbouring LDS arrays alone.
        const uint32_t v = lane < PARTS ? part[lane] : 0u;
        const uint32_t incl = wave_inclusive_sum(v, lane);
        if (lane < PARTS) part[lane] = incl - v;
        if (lane == 63 && tile == a.tiles - 1) *a.n_ranks = before + incl;
    }
    __syncthreads();
    if (BUCKETS) {
        uint32_t lstart = bincl - bcount;
        for (int w = 0; w < wv; w++) lstart += bwt[w];
        bdst[tid] = out_b - lstart;
        bc[tid] = lstart;
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int64_t p = t0 + r * THREADS;
        const bool h = (heads[r] >> lane) & 1;
        // starts up to and including this position, less one
        const uint32_t rk = before + part[r * WAVES + wv] + (uint32_t)__popcll(heads[r] & below) + h - 1;
        if (p <= last) {
            if (BUCKETS)
                sp[bc[row[r] >> a.bshift] + loc[r]] = ((uint64_t)row[r] << 32) | rk;
            else
                a.rank[row[r]] = rk;
            if (h) a.first[rk] = row[r];
        }
    }
    if (BUCKETS) {
        __syncthreads();
        const int len = (int)(a.n - tile * TILE < TILE ? a.n - tile * TILE : TILE);
#pragma unroll
        for (int r = 0; r < KPT; r++) {
            const int i = r * THREADS + tid;
            const uint64_t v = sp[i];
            // (slots past the tile's end hold no pair: the mask keeps their
            // lookup inside bdst)
            const uint32_t pos = bdst[((uint32_t)(v >> 32) >> a.bshift) & 0xff] + (uint32_t)i;
            if (i < len) a.pairs[pos] = v;
        }
    }
}

// rank[row] for the rows of bucket blockIdx.x (2^bshift rows, at most
// 2^15): its pairs, in any order, fill an LDS window of the bucket's ranks,
// which then goes out in whole lines.
constexpr int WINDOW = 1 << 15;

__global__ __launch_bounds__(512) void rows_from_buckets_kernel(const uint64_t *__restrict__ pairs, int64_t n,
                                                                int bshift, uint32_t *__restrict__ rank) {
    __shared__ uint32_t win[WINDOW];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x << bshift;
    const int64_t left = n - r0;
    const int cnt = (int)(left < ((int64_t)1 << bshift) ? left : ((int64_t)1 << bshift));
    const uint64_t *src = pairs + r0;   // the bucket's pairs: as many as its rows, after those of lower rows
    for (int i0 = 0; i0 < cnt; i0 += 8 * 512) {
        uint64_t v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int i = i0 + k * 512 + tid;
            v[k] = src[i < cnt ? i : cnt - 1];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int i = i0 + k * 512 + tid;
            if (i < cnt) win[(uint32_t)((v[k] >> 32) - (uint64_t)r0) & (WINDOW - 1)] = (uint32_t)v[k];
        }
    }
    __syncthreads();
    for (int i = tid; i < cnt; i += 512) rank[r0 + i] = win[i];
}

// Exclusive scan of 1024 values per block (4 per thread); each block's total
// to sums[blockIdx.x] if sums is given.
__global__ __launch_bounds__(256) void scan_block_kernel(const uint32_t *in, uint32_t *out,
                                                         uint32_t *sums, int64_t n) {
    __shared__ uint32_t wave_tot[4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * 1024 + tid * 4;
    uint32_t v[4], t = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i] = base + i < n ? in[base + i] : 0;
        t += v[i];
    }
    const uint32_t inc = wave_inclusive_sum(t, lane);
    if (lane == 63) wave_tot[wv] = inc;
    __syncthreads();
    uint32_t prefix = 0;
    for (int k = 0; k < wv; k++) prefix += wave_tot[k];
    uint32_t run = prefix + inc - t;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (sums && tid == 255) sums[blockIdx.x] = prefix + inc;
}

__global__ __launch_bounds__(256) void add_offsets_kernel(uint32_t *out, const uint32_t *sums, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / 1024];
}

__global__ __launch_bounds__(256) void iota_kernel(uint32_t *out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------
// Ranking by hashing, for inputs whose distinct words are few next to the
// rows (join and grouping keys usually are): the words are deduplicated
// first and only the distinct ones are ordered.
//  1. Each workgroup takes 1024 consecutive rows and deduplicates them in
//     an LDS table (lowest row of each word by LDS atomics), then enters
//     each of its distinct words once into a device table of `cap` slots
//     (open addressing, linear probing, compare-and-swap from EMPTY), takes
//     its lowest row into the slot's first-row word (atomicMin, skipped
//     when the word's first row is already lower: the device-wide atomics
//     on a frequent word's slot would otherwise queue one per workgroup)
//     and writes every row's slot.  Its smallest and largest word go to an
//     array of their own, reduced by one small launch.
//  2. The table's occupied slots are counted into `nb` buckets of equal
//     width between the smallest and the largest word, the counts scanned,
//     and the words scattered by bucket; a word's rank is its bucket's
//     start plus the words of its bucket below it (a bucket holds a few
//     words when the words spread over their range; one of more than
//     BIG_BUCKET words stops this path).
//  3. Each row's rank is its slot's.
// Atomics decide only where a word is stored, never a result: ranks (the
// order of distinct words) and first rows (minima) are the same whatever
// the interleaving.  When a probe runs past HPROBES slots (more distinct
// words than the table takes) or a bucket is too large, the call ranks
// again by the radix sort.  The word EMPTY, which marks free slots, is
// kept apart in slot `cap` (present when its first row is set).
// ---------------------------------------------------------------------
constexpr int HTHREADS = 256;
constexpr int HROWS = 1024;        // rows per workgroup of the insert pass
constexpr int HSLOTS = 2048;       // its LDS table (twice its rows: never full)
constexpr int HPROBES = 128;       // longest probe into the device table
constexpr int BIG_BUCKET = 64;     // largest bucket ranked in place
constexpr unsigned long long EMPTY = ~0ull;
constexpr int64_t HASH_FROM = 1 << 16;   // rows from which ranking hashes first

struct HashState {
    unsigned long long kmin, kmax;   // smallest and largest word
    unsigned flags;                  // nonzero: this path cannot finish
};

struct HashArgs {
    const uint64_t *words;
    int64_t n;
    unsigned long long *key;   // [cap]
    uint32_t *first;           // [cap + 2]: lowest row per slot; slot cap is EMPTY's
    uint32_t *slot;            // [n]
    uint32_t *bcount;          // [nb + 1]: bucket counts, scanned in place to starts
    uint32_t *bpos;            // [cap + 1]: each word's place in its bucket
    uint64_t *bkeys;           // [cap + 1]: the words by bucket
    uint32_t *rank_of;         // [cap + 2]
    unsigned long long *wg_range;   // [2 * workgroups of the insert pass]: smallest, largest word
    HashState *st;
    uint32_t cap, nb;
    int nb_log;
    uint32_t *rank, *first_out;
};

// MurmurHash3's 64-bit finalizer (Appleby): spreads clustered words over
// the table.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

__global__ __launch_bounds__(256) void hash_clear_kernel(HashArgs a) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)a.cap + 2; i += stride) {
        if (i < a.cap) a.key[i] = EMPTY;
        a.first[i] = ~0u;
        if (i <= a.nb) a.bcount[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.st->kmin = EMPTY;
        a.st->kmax = 0;
        a.st->flags = 0;
    }
}

__global__ __launch_bounds__(HTHREADS) void hash_insert_kernel(HashArgs a) {
    __shared__ unsigned long long lkey[HSLOTS];
    // Per local slot (HSLOTS: the word EMPTY's): its lowest row, replaced
    // by its device slot once the word is entered.
    __shared__ uint32_t lmin[HSLOTS + 1];
    __shared__ uint32_t act[HROWS + 1];     // the occupied local slots
    __shared__ unsigned long long wlo[HTHREADS / 64], whi[HTHREADS / 64];
    __shared__ unsigned nact, wg_fail;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * HROWS;
    for (int i = tid; i < HSLOTS; i += HTHREADS) {
        lkey[i] = EMPTY;
        lmin[i] = ~0u;
    }
    if (tid == 0) {
        lmin[HSLOTS] = ~0u;
        nact = 0;
        wg_fail = 0;
    }
    __syncthreads();
    constexpr int PER = HROWS / HTHREADS;
    int ls[PER];
    unsigned long long words[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {   // every load in flight at once (row clamped)
        const int64_t row = r0 + k * HTHREADS + tid;
        words[k] = a.words[row < a.n ? row : a.n - 1];
    }
    unsigned long long lo = EMPTY, hi = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int64_t row = r0 + k * HTHREADS + tid;
        ls[k] = -1;
        if (row >= a.n) continue;
        const unsigned long long w = words[k];
        lo = w < lo ? w : lo;
        hi = w > hi ? w : hi;
        int h = HSLOTS;
        if (w != EMPTY) {
            h = (int)(mix64(w) & (HSLOTS - 1));
            for (;;) {   // ends: the table has twice the workgroup's rows
                unsigned long long cur = __hip_atomic_load(&lkey[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (cur == EMPTY) {
                    cur = atomicCAS(&lkey[h], EMPTY, w);
                    if (cur == EMPTY) break;
                }
                if (cur == w) break;
                h = (h + 1) & (HSLOTS - 1);
            }
        }
        ls[k] = h;
        if (__hip_atomic_load(&lmin[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > (uint32_t)row)
            atomicMin(&lmin[h], (uint32_t)row);
    }
    // The workgroup's smallest and largest word (lanes without rows hold
    // EMPTY and 0, which change neither).
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    if (lane == 0) {
        wlo[wave] = lo;
        whi[wave] = hi;
    }
    __syncthreads();
    // The occupied local slots, packed (in any order).
    for (int base = 0; base <= HSLOTS; base += HTHREADS) {
        const int i = base + tid;
        const bool occupied = i <= HSLOTS && lmin[i] != ~0u;
        const unsigned long long mask = __ballot(occupied);
        if (mask) {
            const int leader = __ffsll((long long)mask) - 1;
            unsigned at = 0;
            if (lane == leader) at = atomicAdd(&nact, (unsigned)__popcll(mask));
            at = __shfl(at, leader) + (unsigned)__popcll(mask & ((1ull << lane) - 1));
            if (occupied) act[at] = (uint32_t)i;
        }
    }
    __syncthreads();
    // Each distinct word of the workgroup into the device table.  A plain
    // load of a slot may return an older value from a cache: EMPTY (the
    // compare-and-swap then reads the slot as it is) or the word the slot
    // holds now (a slot changes once).  A first-row word only decreases,
    // so a value read earlier is never below the present one: when it is
    // not above this workgroup's row, neither is the present one.
    const unsigned na = nact;
    for (unsigned j = tid; j < na; j += HTHREADS) {
        const uint32_t i = act[j];
        const uint32_t mn = lmin[i];
        uint32_t g = a.cap + 1;   // no slot: the path fails, and slot cap + 1 is never read for a result
        if (i == HSLOTS) {
            g = a.cap;
        } else {
            const unsigned long long w = lkey[i];
            uint32_t h = (uint32_t)(mix64(w) & (a.cap - 1));
            for (int p = 0; p < HPROBES; p++) {
                unsigned long long cur = a.key[h];
                if (cur == EMPTY) cur = atomicCAS(&a.key[h], EMPTY, w);
                if (cur == EMPTY || cur == w) {
                    g = h;
                    break;
                }
                h = (h + 1) & (a.cap - 1);
            }
            if (g == a.cap + 1) wg_fail = 1;
        }
        if (g <= a.cap && __hip_atomic_load(&a.first[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > mn)
            atomicMin(&a.first[g], mn);
        lmin[i] = g;
    }
    __syncthreads();
    if (tid == 0) {
        if (wg_fail) atomicOr(&a.st->flags, 1u);
        unsigned long long l = wlo[0], h = whi[0];
        for (int v = 1; v < HTHREADS / 64; v++) {
            l = wlo[v] < l ? wlo[v] : l;
            h = whi[v] > h ? whi[v] : h;
        }
        a.wg_range[2 * blockIdx.x] = l;
        a.wg_range[2 * blockIdx.x + 1] = h;
    }
#pragma unroll
    for (int k = 0; k < PER; k++)
        if (ls[k] >= 0) a.slot[r0 + k * HTHREADS + tid] = lmin[ls[k]];
}

// The smallest and largest word over the insert pass's workgroups (one
// workgroup of 1024 threads).
__global__ __launch_bounds__(1024) void hash_range_kernel(HashArgs a, int groups) {
    __shared__ unsigned long long slo[16], shi[16];
    unsigned long long lo = EMPTY, hi = 0;
    for (int i = threadIdx.x; i < groups; i += 1024) {
        const unsigned long long l = a.wg_range[2 * i], h = a.wg_range[2 * i + 1];
        lo = l < lo ? l : lo;
        hi = h > hi ? h : hi;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
        lo = l2 < lo ? l2 : lo;
        hi = h2 > hi ? h2 : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        slo[threadIdx.x >> 6] = lo;
        shi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < 16; v++) {   // lo, hi: wave 0's
            lo = slo[v] < lo ? slo[v] : lo;
            hi = shi[v] > hi ? shi[v] : hi;
        }
        a.st->kmin = lo;
        a.st->kmax = hi;
    }
}

// Whether table slot g (0 .. cap) holds a word, and which.
__device__ __forceinline__ bool slot_word(const HashArgs &a, int64_t g, uint64_t &w) {
    if (g == a.cap) {
        w = EMPTY;
        return a.first[g] != ~0u;
    }
    w = a.key[g];
    return w != EMPTY;
}

// The bucket of a word: nb buckets of equal width from kmin to kmax.
__device__ __forceinline__ uint32_t bucket_of(uint64_t w, const HashState &st, int nb_log) {
    const uint64_t range = st.kmax - st.kmin;
    const int bits = range ? 64 - __clzll((long long)range) : 0;
    const int shift = bits > nb_log ? bits - nb_log : 0;
    return (uint32_t)((w - st.kmin) >> shift);
}

__global__ __launch_bounds__(256) void bucket_count_kernel(HashArgs a) {
    const HashState st = *a.st;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g <= (int64_t)a.cap; g += stride) {
        uint64_t w;
        if (slot_word(a, g, w)) a.bpos[g] = atomicAdd(&a.bcount[bucket_of(w, st, a.nb_log)], 1u);
    }
}

__global__ __launch_bounds__(256) void bucket_scatter_kernel(HashArgs a) {
    const HashState st = *a.st;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g <= (int64_t)a.cap; g += stride) {
        uint64_t w;
        if (slot_word(a, g, w)) a.bkeys[a.bcount[bucket_of(w, st, a.nb_log)] + a.bpos[g]] = w;
    }
}

__global__ __launch_bounds__(256) void bucket_rank_kernel(HashArgs a) {
    const HashState st = *a.st;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g <= (int64_t)a.cap; g += stride) {
        uint64_t w;
        if (!slot_word(a, g, w)) continue;
        const uint32_t b = bucket_of(w, st, a.nb_log);
        const uint32_t lo = a.bcount[b], hi = a.bcount[b + 1];
        if (hi - lo > (uint32_t)BIG_BUCKET) {
            a.st->flags = 2u;   // any nonzero value stops the path
            continue;
        }
        uint32_t below = 0;
        for (uint32_t j = lo; j < hi; j++) below += a.bkeys[j] < w ? 1u : 0u;
        a.rank_of[g] = lo + below;
        a.first_out[lo + below] = a.first[g];
    }
}

__global__ __launch_bounds__(256) void rank_rows_kernel(HashArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n) a.rank[i] = a.rank_of[a.slot[i]];
}

size_t round_up(size_t x) { return (x + 255) & ~(size_t)255; }

size_t scan_temp_bytes(size_t n) {
    size_t bytes = 0;
    while (n > 1024) {
        n = (n + 1023) / 1024;
        bytes += round_up(n * 4);
    }
    return bytes + 256;
}

int64_t tiles_of(size_t n) { return (int64_t)((n + TILE - 1) / TILE); }


}  // namespace

// Exclusive scan of n u32 values (device), in place allowed, with `tmp` of
// scan_temp_bytes(n) bytes.
ef_status scan_u32(const uint32_t *in, uint32_t *out, size_t n, void *tmp, hipStream_t s) {
    if (n == 0) return EF_OK;
    const int64_t blocks = (int64_t)((n + 1023) / 1024);
    uint32_t *sums = blocks > 1 ? static_cast<uint32_t *>(tmp) : nullptr;
    hipLaunchKernelGGL(scan_block_kernel, dim3((unsigned)blocks), dim3(256), 0, s, in, out, sums, (int64_t)n);
    if (blocks > 1) {
        char *next = static_cast<char *>(tmp) + round_up(blocks * 4);
        EF_TRY(scan_u32(sums, sums, (size_t)blocks, next, s));
        hipLaunchKernelGGL(add_offsets_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, out, sums, (int64_t)n);
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

size_t scan_scratch_bytes(size_t n) { return scan_temp_bytes(n); }

namespace {

size_t radix_bytes(size_t n) {
    const int64_t tiles = tiles_of(n);
    return round_up(n * 8) * 2 + round_up(n * 4) * 2 + round_up((size_t)tiles * 2048 * 4) + 2048 * 4 +
           2 * round_up((size_t)tiles * 257 * 4) + SMALL_WORDS * 4;
}

// Sizes of the hashing path for n rows: a device table of cap >= n / 2
// slots, cap / 8 buckets.
struct HashLayout {
    uint32_t cap, nb;
    int nb_log;
};

HashLayout hash_layout(size_t n) {
    HashLayout L;
    size_t cap = HSLOTS;
    while (cap < (n + 1) / 2) cap <<= 1;
    L.cap = (uint32_t)cap;
    L.nb = L.cap / 8 > 256 ? L.cap / 8 : 256;
    L.nb_log = 0;
    while ((1u << L.nb_log) < L.nb) L.nb_log++;
    return L;
}

// Carves the hashing path's arrays from `scratch` (a null scratch gives
// the bytes needed, in `bytes`).
HashArgs hash_carve(void *scratch, size_t n, const HashLayout &L, size_t &bytes) {
    HashArgs a{};
    char *p = static_cast<char *>(scratch);
    size_t at = 0;
    auto take = [&](size_t b) {
        char *q = p ? p + at : nullptr;
        at += round_up(b);
        return q;
    };
    a.key = reinterpret_cast<unsigned long long *>(take((size_t)L.cap * 8));
    a.first = reinterpret_cast<uint32_t *>(take(((size_t)L.cap + 2) * 4));
    a.slot = reinterpret_cast<uint32_t *>(take(n * 4));
    a.bcount = reinterpret_cast<uint32_t *>(take(((size_t)L.nb + 1) * 4 + scan_temp_bytes((size_t)L.nb + 1)));
    a.bpos = reinterpret_cast<uint32_t *>(take(((size_t)L.cap + 1) * 4));
    a.bkeys = reinterpret_cast<uint64_t *>(take(((size_t)L.cap + 1) * 8));
    a.rank_of = reinterpret_cast<uint32_t *>(take(((size_t)L.cap + 2) * 4));
    a.wg_range = reinterpret_cast<unsigned long long *>(take(((n + HROWS - 1) / HROWS) * 16));
    a.st = reinterpret_cast<HashState *>(take(sizeof(HashState)));
    a.n = (int64_t)n;
    a.cap = L.cap;
    a.nb = L.nb;
    a.nb_log = L.nb_log;
    bytes = at;
    return a;
}

}  // namespace

size_t rank_scratch_bytes(size_t n) {
    size_t bytes = radix_bytes(n);
    if ((int64_t)n >= HASH_FROM) {
        size_t hb = 0;
        hash_carve(nullptr, n, hash_layout(n), hb);
        bytes = hb > bytes ? hb : bytes;
    }
    return bytes;
}

namespace {

// Scratch layout shared by ranking and sorting.
struct RadixScratch {
    uint64_t *wa, *wb;
    uint32_t *ra, *rb, *partial, *counts, *tile_counts, *offsets;
    unsigned *small;
    int64_t tiles;
};

RadixScratch carve(void *scratch, size_t n) {
    RadixScratch r;
    r.tiles = tiles_of(n);
    char *p = static_cast<char *>(scratch);
    r.wa = reinterpret_cast<uint64_t *>(p); p += round_up(n * 8);
    r.wb = reinterpret_cast<uint64_t *>(p); p += round_up(n * 8);
    r.ra = reinterpret_cast<uint32_t *>(p); p += round_up(n * 4);
    r.rb = reinterpret_cast<uint32_t *>(p); p += round_up(n * 4);
    r.partial = reinterpret_cast<uint32_t *>(p); p += round_up((size_t)r.tiles * 2048 * 4);
    r.counts = reinterpret_cast<uint32_t *>(p); p += 2048 * 4;
    r.tile_counts = reinterpret_cast<uint32_t *>(p); p += round_up((size_t)r.tiles * 257 * 4);
    r.offsets = reinterpret_cast<uint32_t *>(p); p += round_up((size_t)r.tiles * 257 * 4);
    r.small = reinterpret_cast<unsigned *>(p);
    return r;
}

// The stable LSD passes over the bytes some word varies in
// (discriminate.rs:113-156).  The last pass writes its rows to r_final and
// its words to w_final when those are given (and differ from the inputs).
// On return (w_cur, r_cur) hold the sorted pairs and `passes` the number
// of passes made; r_cur is null when no byte varies (the input order
// stands).
ef_status radix_passes(const uint64_t *words, const uint32_t *rows_in, size_t n, const RadixScratch &x,
                       uint64_t *w_final, uint32_t *r_final, const uint64_t *&w_cur,
                       const uint32_t *&r_cur, int &passes, hipStream_t s) {
    const HistArgs h{words, (int64_t)n, x.tiles, x.partial, x.small, x.counts};
    hipLaunchKernelGGL(byte_histograms_kernel, dim3((unsigned)x.tiles), dim3(THREADS), 0, s, h);
    hipLaunchKernelGGL(sum_histograms_kernel, dim3(HIST_CHUNKS, 8), dim3(THREADS), 0, s, x.partial, (int)x.tiles,
                       x.counts);
    EF_HIP(hipGetLastError());
    uint32_t counts[2048];
    EF_HIP(hipMemcpyAsync(counts, x.counts, sizeof counts, hipMemcpyDeviceToHost, s));
    EF_HIP(hipStreamSynchronize(s));
    // A byte varies unless one of its values holds every word.
    bool varies[8];
    int last = -1;
    for (int b = 0; b < 8; b++) {
        varies[b] = true;
        for (int d = 0; d < 256; d++)
            if (counts[b * 256 + d] == n) varies[b] = false;
        if (varies[b]) last = b;
    }
    const bool direct = r_final && r_final != rows_in && (!w_final || w_final != words);
    w_cur = words;
    r_cur = rows_in;
    passes = 0;
    bool in_a = false;
    for (int b = 0; b <= last; b++) {
        if (!varies[b]) continue;
        uint64_t *w_next = in_a ? x.wb : x.wa;
        uint32_t *r_next = in_a ? x.rb : x.ra;
        if (b == last && direct) {
            r_next = r_final;
            if (w_final) w_next = w_final;
        }
        if (passes == 0) {
            // The words in input order: the histogram read counted its tiles.
            hipLaunchKernelGGL(tile_offsets_kernel, dim3(256), dim3(THREADS), 0, s, x.partial + 256 * b,
                               x.counts + 256 * b, x.tiles, x.offsets, (int64_t)n, (int64_t)0, (int64_t)1,
                               (int64_t)2048);
        } else {
            hipLaunchKernelGGL(tile_counts_kernel<uint64_t>, dim3((unsigned)x.tiles), dim3(THREADS), 0, s, w_cur,
                               (int64_t)n, 8 * b, x.tile_counts, x.tiles);
            hipLaunchKernelGGL(tile_offsets_kernel, dim3(256), dim3(THREADS), 0, s, x.tile_counts,
                               x.counts + 256 * b, x.tiles, x.offsets, (int64_t)n, (int64_t)0, x.tiles, (int64_t)1);
        }
        const SweepArgs sa{w_cur, r_cur, w_next, r_next, (int64_t)n, 8 * b, x.offsets, x.tiles};
        hipLaunchKernelGGL(sweep_kernel, dim3((unsigned)x.tiles), dim3(THREADS), 0, s, sa);
        w_cur = w_next;
        r_cur = r_next;
        in_a = !in_a;
        passes++;
    }
    if (passes == 0) r_cur = nullptr;
    EF_HIP(hipGetLastError());
    return EF_OK;
}


}  // namespace

ef_status sort_u64(const uint64_t *words, const uint32_t *rows_in, size_t n, uint64_t *w_out, uint32_t *r_out,
                   void *scratch, hipStream_t s) {
    if (n == 0) return EF_OK;
    if (n > 0xffffffffULL) {
        set_error("sorting takes at most 2^32 - 1 rows");
        return EF_INVALID;
    }
    const RadixScratch x = carve(scratch, n);
    const uint64_t *w_cur;
    const uint32_t *r_cur;
    int passes = 0;
    EF_TRY(radix_passes(words, rows_in, n, x, w_out, r_out, w_cur, r_cur, passes, s));
    if (w_out && w_cur != w_out) EF_HIP(hipMemcpyAsync(w_out, w_cur, n * 8, hipMemcpyDeviceToDevice, s));
    if (r_cur) {
        if (r_cur != r_out) EF_HIP(hipMemcpyAsync(r_out, r_cur, n * 4, hipMemcpyDeviceToDevice, s));
    } else if (rows_in) {
        if (rows_in != r_out) EF_HIP(hipMemcpyAsync(r_out, rows_in, n * 4, hipMemcpyDeviceToDevice, s));
    } else {
        hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, r_out, (int64_t)n);
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

namespace {

// Ranking by hashing (the kernels above); `done` false when this path
// cannot finish for these words (the caller ranks them by sorting).
ef_status rank_by_hash(const uint64_t *words, size_t n, uint32_t *rank, uint32_t *first, size_t *n_ranks,
                       void *scratch, bool &done, hipStream_t s) {
    done = false;
    const HashLayout L = hash_layout(n);
    size_t bytes = 0;
    HashArgs a = hash_carve(scratch, n, L, bytes);
    a.words = words;
    a.rank = rank;
    a.first_out = first;
    const unsigned slot_blocks = blocks_for((size_t)L.cap + 2, 256) < 4096 ? blocks_for((size_t)L.cap + 2, 256) : 4096;
    hipLaunchKernelGGL(hash_clear_kernel, dim3(slot_blocks), dim3(256), 0, s, a);
    const unsigned groups = (unsigned)((n + HROWS - 1) / HROWS);
    hipLaunchKernelGGL(hash_insert_kernel, dim3(groups), dim3(HTHREADS), 0, s, a);
    hipLaunchKernelGGL(hash_range_kernel, dim3(1), dim3(1024), 0, s, a, (int)groups);
    hipLaunchKernelGGL(bucket_count_kernel, dim3(slot_blocks), dim3(256), 0, s, a);
    EF_TRY(scan_u32(a.bcount, a.bcount, (size_t)L.nb + 1, a.bcount + L.nb + 1, s));
    hipLaunchKernelGGL(bucket_scatter_kernel, dim3(slot_blocks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(bucket_rank_kernel, dim3(slot_blocks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(rank_rows_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, a);
    EF_HIP(hipGetLastError());
    struct {
        HashState st;
        uint32_t total;
    } h;
    EF_HIP(hipMemcpyAsync(&h.st, a.st, sizeof h.st, hipMemcpyDeviceToHost, s));
    EF_HIP(hipMemcpyAsync(&h.total, a.bcount + L.nb, 4, hipMemcpyDeviceToHost, s));
    EF_HIP(hipStreamSynchronize(s));
    if (h.st.flags) return EF_OK;
    if (n_ranks) *n_ranks = h.total;
    done = true;
    return EF_OK;
}

}  // namespace

ef_status rank_u64(const uint64_t *words, size_t n, uint32_t *rank, uint32_t *first,
                   size_t *n_ranks, void *scratch, hipStream_t s) {
    if (n_ranks) *n_ranks = 0;
    if (n == 0) return EF_OK;
    if (n > 0xffffffffULL) {
        set_error("ranking takes at most 2^32 - 1 rows");
        return EF_INVALID;
    }
    DeviceBuffer own;
    if (!scratch) {
        EF_TRY(own.alloc(rank_scratch_bytes(n)));
        scratch = own.ptr;
    }
    if ((int64_t)n >= HASH_FROM) {
        bool done = false;
        EF_TRY(rank_by_hash(words, n, rank, first, n_ranks, scratch, done, s));
        if (done) return EF_OK;
    }
    const RadixScratch x = carve(scratch, n);
    const uint64_t *w_cur;
    const uint32_t *r_cur;
    int passes = 0;
    EF_TRY(radix_passes(words, nullptr, n, x, nullptr, nullptr, w_cur, r_cur, passes, s));
    if (!r_cur) {
        // Every word is the same: one rank, first row 0.
        EF_HIP(hipMemsetAsync(rank, 0, n * 4, s));
        const uint32_t zero = 0;
        EF_HIP(hipMemcpyAsync(first, &zero, 4, hipMemcpyHostToDevice, s));
        EF_HIP(hipStreamSynchronize(s));
        if (n_ranks) *n_ranks = 1;
        return EF_OK;
    }
    // Rows in buckets of 2^bshift <= WINDOW rows, at most 256 buckets.
    int bshift = 0;
    while (((size_t)256 << bshift) < n) bshift++;
    const bool buckets = n >= ((size_t)1 << 16) && ((size_t)1 << bshift) <= (size_t)WINDOW;
    const int64_t nb = buckets ? (int64_t)((n + ((size_t)1 << bshift) - 1) >> bshift) : 0;
    if (buckets)
        hipLaunchKernelGGL(starts_kernel<true>, dim3((unsigned)x.tiles), dim3(THREADS), 0, s, w_cur, r_cur, (int64_t)n,
                           bshift, x.tile_counts, x.tiles);
    else
        hipLaunchKernelGGL(starts_kernel<false>, dim3((unsigned)x.tiles), dim3(THREADS), 0, s, w_cur, r_cur,
                           (int64_t)n, bshift, x.tile_counts, x.tiles);
    // Row 256 (rank starts): its "values below" are none, so workgroup 0 of
    // a launch based at row 256 gives the exclusive prefix over tiles.
    hipLaunchKernelGGL(tile_offsets_kernel, dim3(1), dim3(THREADS), 0, s, x.tile_counts + 256 * x.tiles, x.counts,
                       x.tiles, x.offsets + 256 * x.tiles, (int64_t)n, (int64_t)0, x.tiles, (int64_t)1);
    AssignArgs aa{w_cur, r_cur, (int64_t)n, x.tiles, rank, first, x.offsets, x.small + S_RANKS, bshift, nullptr};
    if (buckets) {
        hipLaunchKernelGGL(tile_offsets_kernel, dim3((unsigned)nb), dim3(THREADS), 0, s, x.tile_counts,
                           (const uint32_t *)nullptr, x.tiles, x.offsets, (int64_t)n, (int64_t)1 << bshift, x.tiles,
                           (int64_t)1);
        aa.pairs = w_cur == x.wa ? x.wb : x.wa;   // the words buffer the last pass did not write
        hipLaunchKernelGGL(assign_kernel<true>, dim3((unsigned)x.tiles), dim3(THREADS), 0, s, aa);
        hipLaunchKernelGGL(rows_from_buckets_kernel, dim3((unsigned)nb), dim3(512), 0, s, aa.pairs, (int64_t)n, bshift,
                           rank);
    } else {
        hipLaunchKernelGGL(assign_kernel<false>, dim3((unsigned)x.tiles), dim3(THREADS), 0, s, aa);
    }
    EF_HIP(hipGetLastError());
    uint32_t ranks = 0;
    EF_HIP(hipMemcpyAsync(&ranks, x.small + S_RANKS, sizeof ranks, hipMemcpyDeviceToHost, s));
    EF_HIP(hipStreamSynchronize(s));
    if (n_ranks) *n_ranks = ranks;
    return EF_OK;
}

}  // namespace ef
