// Radix ranking on gfx950 (crates/einfold-kernels/src/discriminate.rs:
// 103-156, 280-292): a stable least-significant-byte-first sort of
// (word, row) pairs that skips the bytes no word varies in, then each
// row's rank among the distinct words and each rank's first row.
//
// One read of the words builds all eight byte histograms (and clears the
// look-back words); a small launch adds up the workgroups' histograms, from
// which the host learns which bytes vary.  Then each varying byte is one
// single-pass scatter (the decoupled look-back of Merrill and Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016, used
// for a radix pass as in Adinets and Merrill's "Onesweep", 2022).  Each
// tile of 4096 pairs
//   1. ranks its pairs among the tile's pairs with the same byte, in row
//      order: a wave takes 1024 consecutive rows, 64 at a time, finds the
//      lanes that share its byte with eight ballots and counts them in a
//      histogram of its own in LDS (no barrier inside the wave);
//   2. publishes its 256 counts, then walks back over the tiles before it
//      for the counts of each byte value in them: a tile's word holds
//      either its own count or the count of it and every tile before it,
//      and the walk stops at the first of the latter;
//   3. stages its pairs in LDS in sorted order and writes each byte value's
//      run to its place, so that neighbouring lanes write neighbouring
//      words.
// Tiles take their numbers from a counter in the order they start, so every
// tile a walk waits for is already running.  Each wait is bounded: past the
// limit the kernel sets an error flag (the host reports it) and goes on.
// A last launch counts the distinct words of the sorted pairs the same way
// (one count per tile) and writes each row's rank and each rank's first
// row.  All of it is integer arithmetic: exact and deterministic.
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
constexpr int HIST_BLOCKS = 512;            // workgroups of the histogram read
constexpr unsigned SPIN_LIMIT = 1u << 22;   // polls of one look-back word
constexpr unsigned ASSIGN_PASS = 8;         // look-back tag of the ranking launch

constexpr int HIST_CHUNKS = 64;             // workgroups per byte adding the histograms up

// Small device words of a call: the tile counters of the eight byte passes
// and of the ranking launch, the error flag and the number of ranks.
constexpr int SMALL_WORDS = 64;
constexpr int S_ERR = 9, S_RANKS = 10;
constexpr unsigned ERR_WAIT = 1;   // bit of the error flag

// Look-back words are (tag << 32) | count.  In pass p, tag 2p+1 carries a
// tile's own count and 2p+2 the count of the tile and every tile before
// it; 0 (cleared) or another pass's tag means "not written yet".
__device__ __forceinline__ unsigned long long lb_word(unsigned tag, uint32_t count) {
    return ((unsigned long long)tag << 32) | count;
}

__device__ __forceinline__ void lb_store(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long lb_load(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the counts of tiles 0 .. tile-1 in column `col` (256 words per
// tile) of pass `pass`, one thread per column.  Each round reads the words
// of LB_WINDOW tiles at once (nearest first; tile 0 stands in for tiles
// below it, which the walk never reaches: tile 0's word is the last one it
// can need), adds the own-counts down to the first word that is not
// written yet and stops at the first word that covers every earlier tile.
constexpr int LB_WINDOW = 16;

__device__ uint32_t look_back(const unsigned long long *words, int64_t tile, int col, unsigned pass,
                              unsigned *err) {
    const unsigned own = 2 * pass + 1, all = 2 * pass + 2;
    uint32_t sum = 0;
    unsigned spins = 0;
    for (int64_t j = tile - 1; j >= 0;) {
        unsigned long long v[LB_WINDOW];
#pragma unroll
        for (int w = 0; w < LB_WINDOW; w++) v[w] = lb_load(words + (j - w > 0 ? j - w : 0) * 256 + col);
        int used = 0;
        bool done = false, blocked = false;
#pragma unroll
        for (int w = 0; w < LB_WINDOW; w++) {
            if (done || blocked) continue;
            const unsigned tag = (unsigned)(v[w] >> 32);
            if (tag == all || tag == own) {
                sum += (uint32_t)v[w];
                used++;
                done = tag == all;
            } else {
                blocked = true;
            }
        }
        if (done) return sum;
        j -= used;
        if (used == 0) {
            if (++spins > SPIN_LIMIT) {
                atomicOr(err, ERR_WAIT);
                return sum;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return sum;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// The same walk for one column (column 0) by a whole wave: lane l reads
// tile j - l, 64 tiles a round.  Every lane returns the sum.
__device__ uint32_t look_back_wave(const unsigned long long *words, int64_t tile, unsigned pass, unsigned *err,
                                   int lane) {
    const unsigned own = 2 * pass + 1, all = 2 * pass + 2;
    uint32_t sum = 0;
    unsigned spins = 0;
    for (int64_t j = tile - 1; j >= 0;) {
        const unsigned long long v = lb_load(words + (j - lane > 0 ? j - lane : 0) * 256);
        const unsigned tag = (unsigned)(v >> 32);
        const uint64_t covers = __ballot(tag == all);
        const uint64_t waiting = __ballot(tag != all && tag != own);
        const int first_cover = covers ? __builtin_ctzll(covers) : 64;
        const int first_wait = waiting ? __builtin_ctzll(waiting) : 64;
        if (first_cover < first_wait) return sum + wave_sum(lane <= first_cover ? (uint32_t)v : 0u);
        sum += wave_sum(lane < first_wait ? (uint32_t)v : 0u);
        j -= first_wait;
        if (first_wait == 0) {
            if (++spins > SPIN_LIMIT) {
                if (lane == 0) atomicOr(err, ERR_WAIT);
                return sum;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return sum;
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
    uint32_t *partial;          // [gridDim.x][8][256]
    unsigned long long *lb;     // look-back words to clear
    int64_t lb_n;
    unsigned *small;            // SMALL_WORDS words to clear
    uint32_t *counts;           // [8][256], cleared here, summed by the next launch
};

// The eight byte histograms of each workgroup's tiles, in LDS.
__global__ __launch_bounds__(THREADS) void byte_histograms_kernel(HistArgs a) {
    __shared__ uint32_t h[8][256];
    const int tid = threadIdx.x, lane = tid & 63;
#pragma unroll
    for (int b = 0; b < 8; b++) h[b][tid] = 0;
    for (int64_t i = (int64_t)blockIdx.x * THREADS + tid; i < a.lb_n; i += (int64_t)gridDim.x * THREADS)
        a.lb[i] = 0;
    if (blockIdx.x == 0) {
        if (tid < SMALL_WORDS) a.small[tid] = 0;
#pragma unroll
        for (int b = 0; b < 8; b++) a.counts[b * 256 + tid] = 0;
    }
    __syncthreads();
    for (int64_t t = blockIdx.x; t < a.tiles; t += gridDim.x) {
        const int64_t base = t * TILE;
        uint64_t v[KPT];
#pragma unroll
        for (int r = 0; r < KPT; r++) {
            const int64_t i = base + r * THREADS + tid;
            v[r] = a.w[i < a.n ? i : a.n - 1];
        }
#pragma unroll
        for (int r = 0; r < KPT; r++) {
            const bool valid = base + r * THREADS + tid < a.n;
            const uint64_t live = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                const uint32_t d = (uint32_t)(v[r] >> (8 * b)) & 0xff;
                const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
                // One add for the wave when every live lane has the same
                // byte (live lanes are a prefix of the wave, so lane 0 is
                // one of them).
                if (__ballot(valid && d == d0) == live) {
                    if (lane == 0 && live) atomicAdd(&h[b][d0], (uint32_t)__popcll(live));
                } else if (valid) {
                    atomicAdd(&h[b][d], 1u);
                }
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
    uint32_t s = 0;
    for (int k = c; k < blocks; k += HIST_CHUNKS) s += partial[(int64_t)k * 2048 + b * 256 + tid];
    if (s) atomicAdd(&counts[b * 256 + tid], s);
}

struct SweepArgs {
    const uint64_t *w_in;
    const uint32_t *r_in;       // null: rows 0 .. n-1
    uint64_t *w_out;
    uint32_t *r_out;
    int64_t n;
    int shift;
    unsigned pass;
    const uint32_t *counts;     // the 256 counts of this byte's values
    unsigned long long *lb;     // 256 look-back words per tile
    unsigned *ctr;              // this pass's tile counter
    unsigned *err;
};

// One stable pass over the byte at `shift`.
__global__ __launch_bounds__(THREADS) void sweep_kernel(SweepArgs a) {
    // 52224 bytes of LDS: three workgroups per CU.  The tile number and
    // the scan's wave totals live in `dst` until `dst` is written (after
    // two more barriers).
    __shared__ uint64_t sk[TILE];
    __shared__ uint32_t sr[TILE];
    __shared__ uint16_t cnt[WAVES][256];   // per wave: its pairs of each value, then their place (< TILE)
    __shared__ uint32_t dst[256];          // output position of the tile's value-d run, less its LDS start
    uint32_t *const wt0 = dst, *const wt1 = dst + WAVES;
    unsigned *const tile_s = dst + 2 * WAVES;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) *tile_s = atomicAdd(a.ctr, 1u);
    const uint32_t count_d = a.counts[tid];   // words of value tid in the whole pass
#pragma unroll
    for (int k = 0; k < 4; k++) cnt[wv][k * 64 + lane] = 0;
    __syncthreads();
    const int64_t tile = *tile_s;
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

    // 2. Thread d: the tile's count of value d, published at once; the
    // tile's LDS start of each wave's run of d; the pass's first output
    // position of d.
    const int d = tid;
    uint32_t c[WAVES], total = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        c[w] = cnt[w][d];
        total += c[w];
    }
    unsigned long long *mine = a.lb + tile * 256 + d;
    lb_store(mine, lb_word(tile == 0 ? 2 * a.pass + 2 : 2 * a.pass + 1, total));
    const uint32_t incl = wave_inclusive_sum(total, lane);
    const uint32_t incl_d = wave_inclusive_sum(count_d, lane);
    if (lane == 63) {
        wt0[wv] = incl;
        wt1[wv] = incl_d;
    }
    __syncthreads();
    uint32_t start = incl - total, base_d = incl_d - count_d;
    for (int w = 0; w < wv; w++) {
        start += wt0[w];
        base_d += wt1[w];
    }
    const uint32_t tile_start = start;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        cnt[w][d] = (uint16_t)start;
        start += c[w];
    }
    __syncthreads();

    // 3. The pairs to LDS in sorted order, then the walk back for value d.
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const bool valid = base + r * 64 + lane < a.n;
        const uint32_t lp = cnt[wv][(uint32_t)(key[r] >> a.shift) & 0xff] + off[r];
        if (valid) {
            sk[lp] = key[r];
            sr[lp] = row[r];
        }
    }
    const uint32_t before = look_back(a.lb, tile, d, a.pass, a.err);
    if (tile > 0) lb_store(mine, lb_word(2 * a.pass + 2, before + total));
    dst[d] = base_d + before - tile_start;
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

struct AssignArgs {
    const uint64_t *w;          // sorted words
    const uint32_t *rows;       // their rows
    int64_t n;
    int64_t tiles;
    uint32_t *rank, *first;
    unsigned long long *lb;
    unsigned *ctr, *err;
    uint32_t *n_ranks;
};

// Ranks from the sorted pairs: a position starts a rank where its word
// differs from the one before; the rank of a position is the number of
// starts up to it, less one.  A tile's positions are taken 256 at a time
// (round r, wave, lane), so loads are coalesced; the starts are counted
// per wave and round with ballots.
__global__ __launch_bounds__(THREADS) void assign_kernel(AssignArgs a) {
    __shared__ uint32_t part[KPT * WAVES];   // starts per (round, wave), then their exclusive prefix
    __shared__ unsigned tile_s;
    __shared__ uint32_t before_s;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) tile_s = atomicAdd(a.ctr, 1u);
    __syncthreads();
    const int64_t tile = tile_s;
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
    if (lane < KPT) {
        uint64_t h = heads[0];
#pragma unroll
        for (int r = 1; r < KPT; r++) h = lane == r ? heads[r] : h;
        part[lane * WAVES + wv] = (uint32_t)__popcll(h);
    }
    __syncthreads();
    if (wv == 0) {
        const uint32_t v = part[lane];
        const uint32_t incl = wave_inclusive_sum(v, lane);
        part[lane] = incl - v;
        const uint32_t total = __shfl(incl, 63);
        unsigned long long *mine = a.lb + tile * 256;
        if (lane == 0) lb_store(mine, lb_word(tile == 0 ? 2 * ASSIGN_PASS + 2 : 2 * ASSIGN_PASS + 1, total));
        const uint32_t before = look_back_wave(a.lb, tile, ASSIGN_PASS, a.err, lane);
        if (lane == 0) {
            if (tile > 0) lb_store(mine, lb_word(2 * ASSIGN_PASS + 2, before + total));
            before_s = before;
            if (tile == a.tiles - 1) *a.n_ranks = before + total;
        }
    }
    __syncthreads();
    const uint32_t before = before_s;
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int64_t p = t0 + r * THREADS;
        const bool h = (heads[r] >> lane) & 1;
        // starts up to and including this position, less one
        const uint32_t rk = before + part[r * WAVES + wv] + (uint32_t)__popcll(heads[r] & below) + h - 1;
        if (p <= last) {
            a.rank[row[r]] = rk;
            if (h) a.first[rk] = row[r];
        }
    }
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

int hist_blocks_of(int64_t tiles) { return (int)(tiles < HIST_BLOCKS ? (tiles > 0 ? tiles : 1) : HIST_BLOCKS); }

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

size_t rank_scratch_bytes(size_t n) {
    const int64_t tiles = tiles_of(n);
    const size_t hb = (size_t)hist_blocks_of(tiles);
    return round_up(n * 8) * 2 + round_up(n * 4) * 2 + round_up(hb * 2048 * 4) + 2048 * 4 +
           round_up((size_t)tiles * 256 * 8) + SMALL_WORDS * 4;
}

namespace {

// Scratch layout shared by ranking and sorting.
struct RadixScratch {
    uint64_t *wa, *wb;
    uint32_t *ra, *rb, *partial, *counts;
    unsigned long long *lb;
    unsigned *small;
    int64_t tiles, lb_n;
    int hist_blocks;
};

RadixScratch carve(void *scratch, size_t n) {
    RadixScratch r;
    r.tiles = tiles_of(n);
    r.lb_n = r.tiles * 256;
    r.hist_blocks = hist_blocks_of(r.tiles);
    char *p = static_cast<char *>(scratch);
    r.wa = reinterpret_cast<uint64_t *>(p); p += round_up(n * 8);
    r.wb = reinterpret_cast<uint64_t *>(p); p += round_up(n * 8);
    r.ra = reinterpret_cast<uint32_t *>(p); p += round_up(n * 4);
    r.rb = reinterpret_cast<uint32_t *>(p); p += round_up(n * 4);
    r.partial = reinterpret_cast<uint32_t *>(p); p += round_up((size_t)r.hist_blocks * 2048 * 4);
    r.counts = reinterpret_cast<uint32_t *>(p); p += 2048 * 4;
    r.lb = reinterpret_cast<unsigned long long *>(p); p += round_up((size_t)r.lb_n * 8);
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
    const HistArgs h{words, (int64_t)n, x.tiles, x.partial, x.lb, x.lb_n, x.small, x.counts};
    hipLaunchKernelGGL(byte_histograms_kernel, dim3((unsigned)x.hist_blocks), dim3(THREADS), 0, s, h);
    hipLaunchKernelGGL(sum_histograms_kernel, dim3(HIST_CHUNKS, 8), dim3(THREADS), 0, s, x.partial, x.hist_blocks,
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
        const SweepArgs sa{w_cur, r_cur, w_next, r_next, (int64_t)n, 8 * b, (unsigned)b,
                           x.counts + 256 * b, x.lb, x.small + b, x.small + S_ERR};
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

ef_status wait_error(unsigned err) {
    if (!err) return EF_OK;
    set_error("radix ranking: a tile waited past its limit for the tiles before it");
    return EF_DEVICE_ERROR;
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
    unsigned err = 0;
    if (passes > 0) {
        EF_HIP(hipMemcpyAsync(&err, x.small + S_ERR, 4, hipMemcpyDeviceToHost, s));
        EF_HIP(hipStreamSynchronize(s));
    }
    return wait_error(err);
}

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
    const AssignArgs aa{w_cur, r_cur, (int64_t)n, x.tiles, rank, first, x.lb,
                        x.small + ASSIGN_PASS, x.small + S_ERR, x.small + S_RANKS};
    hipLaunchKernelGGL(assign_kernel, dim3((unsigned)x.tiles), dim3(THREADS), 0, s, aa);
    EF_HIP(hipGetLastError());
    unsigned tail[2] = {0, 0};   // error flag, number of ranks
    EF_HIP(hipMemcpyAsync(tail, x.small + S_ERR, sizeof tail, hipMemcpyDeviceToHost, s));
    EF_HIP(hipStreamSynchronize(s));
    EF_TRY(wait_error(tail[0]));
    if (n_ranks) *n_ranks = tail[1];
    return EF_OK;
}

}  // namespace ef
