// Radix ranking on gfx950 (crates/einfold-kernels/src/discriminate.rs:
// 103-156, 280-292): a stable least-significant-byte-first sort of
// (word, row) pairs that skips the bytes no word varies in, then each
// row's rank among the distinct words and each rank's first row.
//
// A pass is three launches: per-tile byte histograms in LDS; one exclusive
// scan of the counts, byte-major then tile (so every tile's rows of a byte
// land after earlier tiles' rows of that byte: stability); and a scatter in
// which each row's place among its tile's rows of its byte comes from wave
// ballots (eight, one per bit of the byte, give the lanes that share it)
// plus per-wave counts in LDS, in row order.  Everything is integer
// arithmetic: the ranks are exact and deterministic.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.h"

namespace ef {

namespace {

constexpr int THREADS = 256;
constexpr int ROUNDS = 16;                    // keys per thread per tile
constexpr int TILE = THREADS * ROUNDS;        // 4096 keys per tile

// OR and AND of all words: one tile of TILE words per block, all of a
// thread's loads issued together (clamped rows: OR and AND don't change
// when a word is counted twice), one atomic pair per block.
__global__ __launch_bounds__(256) void or_and_kernel(const uint64_t *w, int64_t n,
                                                     unsigned long long *any, unsigned long long *all) {
    __shared__ uint64_t red_o[4], red_a[4];
    const int64_t base = (int64_t)blockIdx.x * TILE;
    uint64_t v[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const int64_t i = base + r * THREADS + threadIdx.x;
        v[r] = w[i < n ? i : n - 1];
    }
    uint64_t o = 0, a = ~0ULL;
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        o |= v[r];
        a &= v[r];
    }
    for (int off = 32; off >= 1; off >>= 1) {
        o |= __shfl_xor(o, off);
        a &= __shfl_xor(a, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red_o[threadIdx.x >> 6] = o;
        red_a[threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicOr(any, (unsigned long long)(red_o[0] | red_o[1] | red_o[2] | red_o[3]));
        atomicAnd(all, (unsigned long long)(red_a[0] & red_a[1] & red_a[2] & red_a[3]));
    }
}

__global__ __launch_bounds__(256) void histogram_kernel(const uint64_t *__restrict__ w, int64_t n, int shift,
                                                        uint32_t *__restrict__ counts, int64_t tiles) {
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * TILE;
    // All of the tile's loads first, from clamped rows (unconditional
    // loads issue together instead of one branch and wait each).
    uint64_t v[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const int64_t i = base + r * THREADS + tid;
        v[r] = w[i < n ? i : n - 1];
    }
#pragma unroll
    for (int r = 0; r < ROUNDS; r++)
        if (base + r * THREADS + tid < n) atomicAdd(&h[(v[r] >> shift) & 0xff], 1u);
    __syncthreads();
    counts[(int64_t)tid * tiles + blockIdx.x] = h[tid];
}

// Stable scatter of one pass.  rows_in == nullptr means rows are 0..n-1.
__global__ __launch_bounds__(256) void scatter_kernel(const uint64_t *__restrict__ w_in,
                                                      const uint32_t *__restrict__ rows_in,
                                                      uint64_t *__restrict__ w_out,
                                                      uint32_t *__restrict__ rows_out, int64_t n,
                                                      int shift, const uint32_t *__restrict__ offsets,
                                                      int64_t tiles) {
    __shared__ uint32_t run[256];          // rows of each byte placed so far
    __shared__ uint32_t wave_count[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    run[tid] = offsets[(int64_t)tid * tiles + blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * TILE;
    const uint64_t below = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
    // The tile's words and rows loaded up front, from clamped rows.
    uint64_t words[ROUNDS];
    uint32_t rws[ROUNDS];
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const int64_t i = base + r * THREADS + tid;
        const int64_t ic = i < n ? i : n - 1;
        words[r] = w_in[ic];
        rws[r] = rows_in ? rows_in[ic] : (uint32_t)i;
    }
#pragma unroll
    for (int r = 0; r < ROUNDS; r++) {
        const int64_t i = base + r * THREADS + tid;
        const bool valid = i < n;
        const uint64_t word = valid ? words[r] : 0;
        const uint32_t row = valid ? rws[r] : 0;
        const uint32_t d = (uint32_t)((word >> shift) & 0xff);
        // Lanes of this wave with the same byte (and a row).
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const uint64_t set = __ballot((d >> bit) & 1);
            peers &= ((d >> bit) & 1) ? set : ~set;
        }
        const uint32_t in_wave = (uint32_t)__popcll(peers & below);
#pragma unroll
        for (int k = 0; k < 4; k++) wave_count[k][tid] = 0;
        __syncthreads();
        // The highest lane of each group of peers writes the group's size.
        if (valid && (peers >> lane) == 1ULL) wave_count[wv][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + in_wave;
            for (int k = 0; k < wv; k++) pos += wave_count[k][d];
            w_out[pos] = word;
            rows_out[pos] = row;
        }
        __syncthreads();
        run[tid] += wave_count[0][tid] + wave_count[1][tid] + wave_count[2][tid] + wave_count[3][tid];
        __syncthreads();
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
    uint32_t inc = t;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(inc, off);
        if (lane >= off) inc += o;
    }
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

// Rank flags: 1 where a sorted word differs from the one before.
__global__ __launch_bounds__(256) void flag_kernel(const uint64_t *w, int64_t n, uint32_t *flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = (i == 0 || w[i] != w[i - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void assign_kernel(const uint64_t *w, const uint32_t *rows, const uint32_t *excl,
                                                     int64_t n, uint32_t *rank, uint32_t *first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool head = i == 0 || w[i] != w[i - 1];
    const uint32_t r = excl[i] + (head ? 1u : 0u) - 1u;
    rank[rows[i]] = r;
    if (head) first[r] = rows[i];
}

__global__ void count_kernel(const uint32_t *excl, const uint32_t *flags, int64_t n, uint32_t *total) {
    *total = excl[n - 1] + flags[n - 1];
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
    const size_t tiles = (n + TILE - 1) / TILE;
    const size_t counts = 256 * tiles;
    return round_up(n * 8) * 2 + round_up(n * 4) * 3 + round_up(counts * 4) * 2 +
           scan_temp_bytes(counts > n ? counts : n) + 1024;
}

namespace {

// Scratch layout shared by ranking and sorting.
struct RadixScratch {
    uint64_t *wa, *wb;
    uint32_t *ra, *rb, *flags, *counts, *offs;
    unsigned long long *bits;
    void *tmp;
    size_t counts_n;
    int64_t tiles;
};

RadixScratch carve(void *scratch, size_t n) {
    RadixScratch r;
    r.tiles = (int64_t)((n + TILE - 1) / TILE);
    r.counts_n = 256 * (size_t)r.tiles;
    char *p = static_cast<char *>(scratch);
    r.wa = reinterpret_cast<uint64_t *>(p); p += round_up(n * 8);
    r.wb = reinterpret_cast<uint64_t *>(p); p += round_up(n * 8);
    r.ra = reinterpret_cast<uint32_t *>(p); p += round_up(n * 4);
    r.rb = reinterpret_cast<uint32_t *>(p); p += round_up(n * 4);
    r.flags = reinterpret_cast<uint32_t *>(p); p += round_up(n * 4);
    r.counts = reinterpret_cast<uint32_t *>(p); p += round_up(r.counts_n * 4);
    r.offs = reinterpret_cast<uint32_t *>(p); p += round_up(r.counts_n * 4);
    r.bits = reinterpret_cast<unsigned long long *>(p); p += 256;
    r.tmp = p;
    return r;
}

// The stable LSD passes over the bytes some word varies in
// (discriminate.rs:113-156).  On return (w_cur, r_cur) hold the sorted
// pairs; r_cur is null when no byte varies (the input order stands).
ef_status radix_passes(const uint64_t *words, const uint32_t *rows_in, size_t n, const RadixScratch &x,
                       const uint64_t *&w_cur, const uint32_t *&r_cur, hipStream_t s) {
    const unsigned long long init[2] = {0ULL, ~0ULL};
    EF_HIP(hipMemcpyAsync(x.bits, init, sizeof init, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(or_and_kernel, dim3((unsigned)x.tiles), dim3(256), 0, s, words, (int64_t)n, x.bits, x.bits + 1);
    unsigned long long host_bits[2];
    EF_HIP(hipMemcpyAsync(host_bits, x.bits, sizeof host_bits, hipMemcpyDeviceToHost, s));
    EF_HIP(hipStreamSynchronize(s));
    const uint64_t varies = host_bits[0] ^ host_bits[1];
    w_cur = words;
    r_cur = rows_in;
    bool sorted_any = false, in_a = false;
    for (int shift = 0; shift < 64; shift += 8) {
        if (((varies >> shift) & 0xff) == 0) continue;
        hipLaunchKernelGGL(histogram_kernel, dim3((unsigned)x.tiles), dim3(256), 0, s, w_cur, (int64_t)n, shift,
                           x.counts, x.tiles);
        EF_TRY(scan_u32(x.counts, x.offs, x.counts_n, x.tmp, s));
        uint64_t *w_next = in_a ? x.wb : x.wa;
        uint32_t *r_next = in_a ? x.rb : x.ra;
        hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)x.tiles), dim3(256), 0, s, w_cur, r_cur, w_next, r_next,
                           (int64_t)n, shift, x.offs, x.tiles);
        w_cur = w_next;
        r_cur = r_next;
        in_a = !in_a;
        sorted_any = true;
    }
    if (!sorted_any) r_cur = nullptr;
    EF_HIP(hipGetLastError());
    return EF_OK;
}

__global__ __launch_bounds__(256) void iota_kernel(uint32_t *out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)i;
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
    EF_TRY(radix_passes(words, rows_in, n, x, w_cur, r_cur, s));
    if (w_out) EF_HIP(hipMemcpyAsync(w_out, w_cur, n * 8, hipMemcpyDeviceToDevice, s));
    if (r_cur) {
        EF_HIP(hipMemcpyAsync(r_out, r_cur, n * 4, hipMemcpyDeviceToDevice, s));
    } else if (rows_in) {
        EF_HIP(hipMemcpyAsync(r_out, rows_in, n * 4, hipMemcpyDeviceToDevice, s));
    } else {
        hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, r_out, (int64_t)n);
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
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
    EF_TRY(radix_passes(words, nullptr, n, x, w_cur, r_cur, s));
    if (!r_cur) {
        // Every word is the same: one rank, first row 0, and the identity
        // order.
        EF_HIP(hipMemsetAsync(rank, 0, n * 4, s));
        const uint32_t zero = 0;
        EF_HIP(hipMemcpyAsync(first, &zero, 4, hipMemcpyHostToDevice, s));
        EF_HIP(hipStreamSynchronize(s));
        if (n_ranks) *n_ranks = 1;
        return EF_OK;
    }
    hipLaunchKernelGGL(flag_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, w_cur, (int64_t)n, x.flags);
    // The row buffer the last pass did not write is free for the scan.
    uint32_t *excl = r_cur == x.ra ? x.rb : x.ra;
    EF_TRY(scan_u32(x.flags, excl, n, x.tmp, s));
    hipLaunchKernelGGL(assign_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, w_cur, r_cur, excl, (int64_t)n,
                       rank, first);
    uint32_t *total = reinterpret_cast<uint32_t *>(x.bits);
    hipLaunchKernelGGL(count_kernel, dim3(1), dim3(1), 0, s, excl, x.flags, (int64_t)n, total);
    uint32_t host_total = 0;
    EF_HIP(hipMemcpyAsync(&host_total, total, 4, hipMemcpyDeviceToHost, s));
    EF_HIP(hipStreamSynchronize(s));
    EF_HIP(hipGetLastError());
    if (n_ranks) *n_ranks = host_total;
    return EF_OK;
}

}  // namespace ef
