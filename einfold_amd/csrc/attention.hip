// Fused attention pipeline on gfx950: the reference's tile algorithm
// (crates/einfold-kernels/src/fused.rs:694-1019) with one score tile per
// wave, an online softmax in registers, and key tiles that no query of a
// workgroup may see never visited; plus backward kernels that recompute the
// score tiles from the saved row max and row sum.
//
// Layout: q [B][Tq][D], k [B][Tk][D], v [B][Tk][DV], out [B][Tq][DV],
// row statistics [B][Tq].  A workgroup is 4 waves; each wave owns 16 rows
// (queries in the forward and dQ kernels, keys in the dK/dV kernel) and
// works on 64-wide tiles of the other side, so a workgroup's tiles are the
// reference's 64 x 64 tile pairs (fused.rs:109-110).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "common.h"
#include "mfma.h"

namespace ef {

namespace {

constexpr int BM = 64;   // rows per workgroup (4 waves x 16)
constexpr int BN = 64;   // columns per tile
constexpr double LOG2E = 1.4426950408889634;   // cast to the precision's scalar
constexpr double LN2 = 0.6931471805599453;

struct Span { int64_t lo, hi; };

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

struct FwdArgs {
    const void *q, *k, *v;
    void *out;
    float *row_max, *row_sum;
    unsigned long long *visited;
    Mask mask;
    int64_t tq, tk;
    double scale;
    int reverse;   // visit query tiles last-first (causal load balance)
};

template <class T> __device__ __forceinline__ T neg_inf() { return -INFINITY; }

// The forward pass for D = DV features, in precision P.
template <int P, int D>
__global__ __launch_bounds__(256) void attn_fwd_kernel(FwdArgs a) {
    typedef typename Input<P>::T In;
    typedef typename Acc<P>::type AccV;
    typedef typename Acc<P>::scalar S;
    typedef typename Elem<P>::T E;
    constexpr int KC = D / 32;           // 32-wide chunks of the features
    constexpr int DB = D / 16;           // 16-wide output blocks
    constexpr int KP = D + 8;            // K tile row pitch (elements)
    constexpr int VP = BN + 8;           // V^T tile row pitch
    constexpr int KPLANE = BN * KP;
    constexpr int VPLANE = D * VP;
    __shared__ __attribute__((aligned(16))) E ks[Elem<P>::planes * KPLANE];
    __shared__ __attribute__((aligned(16))) E vt[Elem<P>::planes * VPLANE];
    __shared__ int64_t red_lo[4], red_hi[4];
    __shared__ int flags[2];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, c = lane & 15;
    const int64_t nqt = (a.tq + BM - 1) / BM;
    const int64_t qt = a.reverse ? nqt - 1 - blockIdx.x : blockIdx.x;
    const int64_t b = blockIdx.y;
    const int64_t q0 = qt * BM;
    const int64_t x = q0 + w * 16 + c;   // this lane's query
    const In *Q = static_cast<const In *>(a.q) + b * a.tq * D;
    const In *K = static_cast<const In *>(a.k) + b * a.tk * D;
    const In *V = static_cast<const In *>(a.v) + b * a.tk * D;

    const Span sp = query_span(a.mask, x);
    // The workgroup's key range: the union of its queries' spans.
    {
        int64_t lo = sp.lo < sp.hi ? sp.lo : INT64_MAX;
        int64_t hi = sp.lo < sp.hi ? sp.hi : 0;
        for (int o = 32; o >= 1; o >>= 1) {
            int64_t l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
            lo = l2 < lo ? l2 : lo;
            hi = h2 > hi ? h2 : hi;
        }
        if (lane == 0) { red_lo[w] = lo; red_hi[w] = hi; }
    }
    __syncthreads();
    int64_t first = red_lo[0], last = red_hi[0];
    for (int i = 1; i < 4; i++) {
        first = red_lo[i] < first ? red_lo[i] : first;
        last = red_hi[i] > last ? red_hi[i] : last;
    }

    // This lane's queries as B fragments of S^T = K Q^T (column = query).
    Frag<P> qf[KC];
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
        In tmp[8];
#pragma unroll
        for (int j = 0; j < 8; j++) tmp[j] = x < a.tq ? Q[x * D + kc * 32 + g * 8 + j] : In(0);
        make_frag<P>(qf[kc], tmp);
    }

    const S scale2 = (S)a.scale * (S)LOG2E;   // scores in log2 units
    S m = neg_inf<S>(), l = 0;
    AccV o[DB];
#pragma unroll
    for (int i = 0; i < DB; i++) o[i] = AccV{0, 0, 0, 0};
    unsigned visited = 0;

    if (first < last) {
        for (int64_t j0 = first - first % BN; j0 < last; j0 += BN) {
            // Whether any query sees this tile, and whether every query
            // sees all of it (then nothing is masked).
            const bool any = sp.lo < sp.hi && sp.lo < j0 + BN && sp.hi > j0;
            const bool full = x >= a.tq || (sp.lo <= j0 && sp.hi >= j0 + BN);
            const int any_wg = __syncthreads_or(any);
            const int full_wg = __syncthreads_and(full);
            if (!any_wg) continue;
            visited++;
            // Stage K [key][d] and V^T [d][key] (zero past the keys).
            for (int idx = tid * 4; idx < BN * D; idx += 256 * 4) {
                const int r = idx / D, col = idx % D;
                const int64_t j = j0 + r;
                In kv[4], vv[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    kv[i] = j < a.tk ? K[j * D + col + i] : In(0);
                    vv[i] = j < a.tk ? V[j * D + col + i] : In(0);
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    tile_store<P>(ks, KPLANE, r * KP + col + i, kv[i]);
                    tile_store<P>(vt, VPLANE, (col + i) * VP + r, vv[i]);
                }
            }
            __syncthreads();
            // S^T block kb: keys j0 + kb*16 + crow(g, i), query x.
            AccV s[4];
#pragma unroll
            for (int kb = 0; kb < 4; kb++) {
                s[kb] = AccV{0, 0, 0, 0};
#pragma unroll
                for (int kc = 0; kc < KC; kc++) {
                    Frag<P> kf;
                    tile_load8<P>(kf, ks, KPLANE, (kb * 16 + c) * KP + kc * 32 + g * 8);
                    mma<P>(s[kb], kf, qf[kc]);
                }
            }
            // Online softmax for query x over this tile.
            S mx = neg_inf<S>();
#pragma unroll
            for (int kb = 0; kb < 4; kb++) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    S t = s[kb][i] * scale2;
                    if (!full_wg) {
                        const int64_t j = j0 + kb * 16 + crow<P>(g, i);
                        if (j < sp.lo || j >= sp.hi) t = neg_inf<S>();
                    }
                    s[kb][i] = t;
                    mx = t > mx ? t : mx;
                }
            }
            mx = wave_max(mx);
            const S mnew = mx > m ? mx : m;
            S p[4][4];
            S rs = 0;
            if (mnew == neg_inf<S>()) {
#pragma unroll
                for (int kb = 0; kb < 4; kb++)
#pragma unroll
                    for (int i = 0; i < 4; i++) p[kb][i] = 0;
            } else {
#pragma unroll
                for (int kb = 0; kb < 4; kb++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        p[kb][i] = exp2(s[kb][i] - mnew);
                        rs += p[kb][i];
                    }
                const S alpha = exp2(m - mnew);   // 0 when m was -inf
                rs = group_sum(rs);
                l = l * alpha + rs;
#pragma unroll
                for (int db = 0; db < DB; db++)
#pragma unroll
                    for (int i = 0; i < 4; i++) o[db][i] *= alpha;
                m = mnew;
            }
            // O^T += V^T P^T: P^T's 32-key chunks from two stacked S blocks.
#pragma unroll
            for (int kc = 0; kc < 2; kc++) {
                Frag<P> pf;
                S tmp[8];
#pragma unroll
                for (int i = 0; i < 4; i++) { tmp[i] = p[2 * kc][i]; tmp[i + 4] = p[2 * kc + 1][i]; }
                make_frag<P>(pf, tmp);
#pragma unroll
                for (int db = 0; db < DB; db++) {
                    Frag<P> vf;
                    tile_load_cperm<P>(vf, vt, VPLANE, (db * 16 + c) * VP + kc * 32, g);
                    mma<P>(o[db], vf, pf);
                }
            }
            __syncthreads();
        }
    }
    // out[x][d] for d = db*16 + crow(g, i): the weighted sums over the
    // normalizer (a query with no keys has l = 0 and gets 0/0 = NaN).
    if (x < a.tq) {
        In *O = static_cast<In *>(a.out) + b * a.tq * D + x * D;
#pragma unroll
        for (int db = 0; db < DB; db++)
#pragma unroll
            for (int i = 0; i < 4; i++) O[db * 16 + crow<P>(g, i)] = (In)(o[db][i] / l);
        if (g == 0) {
            a.row_max[b * a.tq + x] = (float)(m * (S)LN2);
            a.row_sum[b * a.tq + x] = (float)l;
        }
    }
    if (tid == 0 && a.visited) atomicAdd(a.visited, (unsigned long long)visited);
}

// Any shape, one thread per (query, block of 8 value columns), keys in
// order: the reference's algorithm with a tile of one row.  Float64 or
// Float32 arithmetic throughout.
template <class T>
__global__ __launch_bounds__(256) void attn_fwd_generic(const T *Q, const T *K, const T *V, T *O,
                                                         float *row_max, float *row_sum, Mask mask,
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
        row_max[b * tq + x] = (float)m;
        row_sum[b * tq + x] = (float)z;
    }
}

// ---------------------------------------------------------------------
// Backward
// ---------------------------------------------------------------------
struct BwdArgs {
    const void *q, *k, *v, *out, *dout;
    const float *row_max, *row_sum;
    float *delta;
    void *dq, *dk, *dv;
    Mask mask;
    int64_t tq, tk;
    double scale;
};

// delta[b][x] = sum_d dout[b][x][d] * out[b][x][d].
template <class In, int D>
__global__ __launch_bounds__(256) void attn_delta_kernel(BwdArgs a, int64_t rows) {
    const int64_t r = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int c = threadIdx.x & 15;
    float s = 0;
    if (r < rows) {
        const In *o = static_cast<const In *>(a.out) + r * D;
        const In *d = static_cast<const In *>(a.dout) + r * D;
        for (int e = c; e < D; e += 16) s += (float)o[e] * (float)d[e];
    }
    // Fixed-order tree over the 16 lanes of the row.
    for (int off = 8; off >= 1; off >>= 1) s += __shfl_xor(s, off, 16);
    if (r < rows && c == 0) a.delta[r] = s;
}

// dK and dV for 64 keys per workgroup (16 per wave), visiting the query
// tiles that may see them.
template <int P, int D>
__global__ __launch_bounds__(256) void attn_bwd_dkdv_kernel(BwdArgs a, int64_t qt_begin, int64_t qt_end) {
    typedef typename Input<P>::T In;
    typedef typename Acc<P>::type AccV;
    typedef typename Acc<P>::scalar S;
    typedef typename Elem<P>::T E;
    constexpr int KC = D / 32, DB = D / 16;
    constexpr int RP = D + 8;    // row-major tile pitch [q][d]
    constexpr int TP = BM + 8;   // transposed tile pitch [d][q]
    constexpr int RPLANE = BM * RP, TPLANE = D * TP;
    __shared__ __attribute__((aligned(16))) E qs[Elem<P>::planes * RPLANE];
    __shared__ __attribute__((aligned(16))) E qtr[Elem<P>::planes * TPLANE];
    __shared__ __attribute__((aligned(16))) E dos[Elem<P>::planes * RPLANE];
    __shared__ __attribute__((aligned(16))) E dotr[Elem<P>::planes * TPLANE];
    __shared__ float sm[BM], sl[BM], sd[BM];
    __shared__ int64_t slo[BM], shi[BM];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, c = lane & 15;
    const int64_t b = blockIdx.y;
    const int64_t k0 = (int64_t)blockIdx.x * BM;
    const int64_t kj = k0 + w * 16 + c;   // this lane's key (as a column)
    const In *Q = static_cast<const In *>(a.q) + b * a.tq * D;
    const In *K = static_cast<const In *>(a.k) + b * a.tk * D;
    const In *V = static_cast<const In *>(a.v) + b * a.tk * D;
    const In *DO = static_cast<const In *>(a.dout) + b * a.tq * D;
    const float *RM = a.row_max + b * a.tq, *RS = a.row_sum + b * a.tq, *DL = a.delta + b * a.tq;

    Frag<P> kf[KC], vf[KC];
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
        In tk8[8], tv8[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            tk8[j] = kj < a.tk ? K[kj * D + kc * 32 + g * 8 + j] : In(0);
            tv8[j] = kj < a.tk ? V[kj * D + kc * 32 + g * 8 + j] : In(0);
        }
        make_frag<P>(kf[kc], tk8);
        make_frag<P>(vf[kc], tv8);
    }
    AccV dk[DB], dv[DB];
#pragma unroll
    for (int i = 0; i < DB; i++) { dk[i] = AccV{0, 0, 0, 0}; dv[i] = AccV{0, 0, 0, 0}; }
    const S scale2 = (S)a.scale * (S)LOG2E;

    // Under a causal mask alone, query x sees key j when j <= x + offset,
    // so no query below k0 - offset sees this workgroup's keys.
    if (!a.mask.allowed && a.mask.causal && a.mask.window <= 0) {
        const int64_t xmin = k0 - a.mask.offset;
        const int64_t first_tile = xmin > 0 ? xmin / BM : 0;
        qt_begin = first_tile > qt_begin ? first_tile : qt_begin;
    }
    for (int64_t qt = qt_begin; qt < qt_end; qt++) {
        const int64_t q0 = qt * BM;
        if (tid < BM) {
            const int64_t x = q0 + tid;
            const Span sp = query_span(a.mask, x);
            slo[tid] = sp.lo;
            shi[tid] = sp.hi;
            sm[tid] = x < a.tq ? RM[x] : 0.f;
            sl[tid] = x < a.tq ? RS[x] : 1.f;
            sd[tid] = x < a.tq ? DL[x] : 0.f;
        }
        __syncthreads();
        // Whether any query of the tile sees any key of this workgroup.
        bool any = false;
        if (tid < BM) {
            const int64_t kend = k0 + BM < a.tk ? k0 + BM : a.tk;
            any = slo[tid] < shi[tid] && slo[tid] < kend && shi[tid] > k0;
        }
        if (!__syncthreads_or(any)) continue;
        for (int idx = tid * 4; idx < BM * D; idx += 256 * 4) {
            const int r = idx / D, col = idx % D;
            const int64_t x = q0 + r;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const In qv = x < a.tq ? Q[x * D + col + i] : In(0);
                const In dv_ = x < a.tq ? DO[x * D + col + i] : In(0);
                tile_store<P>(qs, RPLANE, r * RP + col + i, qv);
                tile_store<P>(qtr, TPLANE, (col + i) * TP + r, qv);
                tile_store<P>(dos, RPLANE, r * RP + col + i, dv_);
                tile_store<P>(dotr, TPLANE, (col + i) * TP + r, dv_);
            }
        }
        __syncthreads();
        // S = Q K^T and dP = dO V^T: blocks qb, lane holds queries
        // qb*16 + crow(g, i) for key kj.
        AccV s[4], dp[4];
#pragma unroll
        for (int qb = 0; qb < 4; qb++) {
            s[qb] = AccV{0, 0, 0, 0};
            dp[qb] = AccV{0, 0, 0, 0};
#pragma unroll
            for (int kc = 0; kc < KC; kc++) {
                Frag<P> af;
                tile_load8<P>(af, qs, RPLANE, (qb * 16 + c) * RP + kc * 32 + g * 8);
                mma<P>(s[qb], af, kf[kc]);
                tile_load8<P>(af, dos, RPLANE, (qb * 16 + c) * RP + kc * 32 + g * 8);
                mma<P>(dp[qb], af, vf[kc]);
            }
        }
        // P = exp(scale s - m) / l and dS = P (dP - delta), masked to 0.
        S p[4][4], ds[4][4];
#pragma unroll
        for (int qb = 0; qb < 4; qb++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = qb * 16 + crow<P>(g, i);
                const bool ok = kj >= slo[r] && kj < shi[r];
                const S pv = ok ? exp2(s[qb][i] * scale2 - (S)sm[r] * (S)LOG2E) / (S)sl[r] : S(0);
                p[qb][i] = pv;
                ds[qb][i] = pv * (dp[qb][i] - (S)sd[r]);
            }
        // dV += P^T dO and dK += dS^T Q over the tile's 64 queries.
#pragma unroll
        for (int kc = 0; kc < 2; kc++) {
            Frag<P> pf, dsf;
            S t1[8], t2[8];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                t1[i] = p[2 * kc][i]; t1[i + 4] = p[2 * kc + 1][i];
                t2[i] = ds[2 * kc][i]; t2[i + 4] = ds[2 * kc + 1][i];
            }
            make_frag<P>(pf, t1);
            make_frag<P>(dsf, t2);
#pragma unroll
            for (int db = 0; db < DB; db++) {
                Frag<P> bf;
                const int row = (db * 16 + c) * TP + kc * 32;
                tile_load_cperm<P>(bf, dotr, TPLANE, row, g);
                mma<P>(dv[db], pf, bf);
                tile_load_cperm<P>(bf, qtr, TPLANE, row, g);
                mma<P>(dk[db], dsf, bf);
            }
        }
        __syncthreads();
    }
    // Lane holds dK/dV[key k0 + w*16 + crow(g, i)][d = db*16 + c].
    In *DK = static_cast<In *>(a.dk) + b * a.tk * D;
    In *DV = static_cast<In *>(a.dv) + b * a.tk * D;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int64_t key = k0 + w * 16 + crow<P>(g, i);
        if (key < a.tk) {
#pragma unroll
            for (int db = 0; db < DB; db++) {
                DK[key * D + db * 16 + c] = (In)(dk[db][i] * (S)a.scale);
                DV[key * D + db * 16 + c] = (In)dv[db][i];
            }
        }
    }
}

// dQ for 64 queries per workgroup (16 per wave), visiting the key tiles
// each may see.
template <int P, int D>
__global__ __launch_bounds__(256) void attn_bwd_dq_kernel(BwdArgs a) {
    typedef typename Input<P>::T In;
    typedef typename Acc<P>::type AccV;
    typedef typename Acc<P>::scalar S;
    typedef typename Elem<P>::T E;
    constexpr int KC = D / 32, DB = D / 16;
    constexpr int RP = D + 8, TP = BN + 8;
    constexpr int RPLANE = BN * RP, TPLANE = D * TP;
    __shared__ __attribute__((aligned(16))) E ks[Elem<P>::planes * RPLANE];
    __shared__ __attribute__((aligned(16))) E ktr[Elem<P>::planes * TPLANE];
    __shared__ __attribute__((aligned(16))) E vs[Elem<P>::planes * RPLANE];
    __shared__ int64_t red_lo[4], red_hi[4];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, c = lane & 15;
    const int64_t b = blockIdx.y;
    const int64_t q0 = (int64_t)blockIdx.x * BM;
    const int64_t x = q0 + w * 16 + c;
    const In *Q = static_cast<const In *>(a.q) + b * a.tq * D;
    const In *K = static_cast<const In *>(a.k) + b * a.tk * D;
    const In *V = static_cast<const In *>(a.v) + b * a.tk * D;
    const In *DO = static_cast<const In *>(a.dout) + b * a.tq * D;

    const Span sp = query_span(a.mask, x);
    {
        int64_t lo = sp.lo < sp.hi ? sp.lo : INT64_MAX;
        int64_t hi = sp.lo < sp.hi ? sp.hi : 0;
        for (int o = 32; o >= 1; o >>= 1) {
            int64_t l2 = __shfl_xor(lo, o), h2 = __shfl_xor(hi, o);
            lo = l2 < lo ? l2 : lo;
            hi = h2 > hi ? h2 : hi;
        }
        if (lane == 0) { red_lo[w] = lo; red_hi[w] = hi; }
    }
    __syncthreads();
    int64_t first = red_lo[0], last = red_hi[0];
    for (int i = 1; i < 4; i++) {
        first = red_lo[i] < first ? red_lo[i] : first;
        last = red_hi[i] > last ? red_hi[i] : last;
    }
    Frag<P> qf[KC], df[KC];
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
        In t1[8], t2[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            t1[j] = x < a.tq ? Q[x * D + kc * 32 + g * 8 + j] : In(0);
            t2[j] = x < a.tq ? DO[x * D + kc * 32 + g * 8 + j] : In(0);
        }
        make_frag<P>(qf[kc], t1);
        make_frag<P>(df[kc], t2);
    }
    const S mrow = x < a.tq ? (S)a.row_max[b * a.tq + x] * (S)LOG2E : S(0);
    const S lrow = x < a.tq ? (S)a.row_sum[b * a.tq + x] : S(1);
    const S drow = x < a.tq ? (S)a.delta[b * a.tq + x] : S(0);
    const S scale2 = (S)a.scale * (S)LOG2E;
    AccV dq[DB];
#pragma unroll
    for (int i = 0; i < DB; i++) dq[i] = AccV{0, 0, 0, 0};

    if (first < last) {
        for (int64_t j0 = first - first % BN; j0 < last; j0 += BN) {
            const bool any = sp.lo < sp.hi && sp.lo < j0 + BN && sp.hi > j0;
            if (!__syncthreads_or(any)) continue;
            for (int idx = tid * 4; idx < BN * D; idx += 256 * 4) {
                const int r = idx / D, col = idx % D;
                const int64_t j = j0 + r;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const In kv = j < a.tk ? K[j * D + col + i] : In(0);
                    const In vv = j < a.tk ? V[j * D + col + i] : In(0);
                    tile_store<P>(ks, RPLANE, r * RP + col + i, kv);
                    tile_store<P>(ktr, TPLANE, (col + i) * TP + r, kv);
                    tile_store<P>(vs, RPLANE, r * RP + col + i, vv);
                }
            }
            __syncthreads();
            // S^T = K Q^T and dP^T = V dO^T: lane holds keys kb*16 + crow(g, i)
            // for query x.
            AccV s[4], dp[4];
#pragma unroll
            for (int kb = 0; kb < 4; kb++) {
                s[kb] = AccV{0, 0, 0, 0};
                dp[kb] = AccV{0, 0, 0, 0};
#pragma unroll
                for (int kc = 0; kc < KC; kc++) {
                    Frag<P> af;
                    tile_load8<P>(af, ks, RPLANE, (kb * 16 + c) * RP + kc * 32 + g * 8);
                    mma<P>(s[kb], af, qf[kc]);
                    tile_load8<P>(af, vs, RPLANE, (kb * 16 + c) * RP + kc * 32 + g * 8);
                    mma<P>(dp[kb], af, df[kc]);
                }
            }
            S ds[4][4];
#pragma unroll
            for (int kb = 0; kb < 4; kb++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int64_t j = j0 + kb * 16 + crow<P>(g, i);
                    const bool ok = j >= sp.lo && j < sp.hi;
                    const S pv = ok ? exp2(s[kb][i] * scale2 - mrow) / lrow : S(0);
                    ds[kb][i] = pv * (dp[kb][i] - drow);
                }
            // dQ^T += K^T dS^T: A = K^T (d x key, from the transposed tile),
            // B = dS^T (key x query) from registers.
#pragma unroll
            for (int kc = 0; kc < 2; kc++) {
                Frag<P> dsf;
                S t[8];
#pragma unroll
                for (int i = 0; i < 4; i++) { t[i] = ds[2 * kc][i]; t[i + 4] = ds[2 * kc + 1][i]; }
                make_frag<P>(dsf, t);
#pragma unroll
                for (int db = 0; db < DB; db++) {
                    Frag<P> kt;
                    tile_load_cperm<P>(kt, ktr, TPLANE, (db * 16 + c) * TP + kc * 32, g);
                    mma<P>(dq[db], kt, dsf);
                }
            }
            __syncthreads();
        }
    }
    // Lane holds dQ^T[d = db*16 + crow(g, i)][query x].
    if (x < a.tq) {
        In *DQ = static_cast<In *>(a.dq) + b * a.tq * D + x * D;
#pragma unroll
        for (int db = 0; db < DB; db++)
#pragma unroll
            for (int i = 0; i < 4; i++) DQ[db * 16 + crow<P>(g, i)] = (In)(dq[db][i] * (S)a.scale);
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

template <int P, int D>
void launch_fwd(const FwdArgs &a, int64_t batches, hipStream_t s) {
    dim3 grid((unsigned)((a.tq + BM - 1) / BM), (unsigned)batches);
    hipLaunchKernelGGL((attn_fwd_kernel<P, D>), grid, dim3(256), 0, s, a);
}

template <int P>
bool dispatch_fwd(const FwdArgs &a, int64_t batches, int64_t d, hipStream_t s) {
    switch (d) {
    case 32: launch_fwd<P, 32>(a, batches, s); return true;
    case 64: launch_fwd<P, 64>(a, batches, s); return true;
    case 128:
        if constexpr (P == PREC_F64) return false;   // LDS: the generic kernel takes it
        else { launch_fwd<P, 128>(a, batches, s); return true; }
    default: return false;
    }
}

template <int P, int D>
void launch_bwd(const BwdArgs &a, int64_t batches, int stage, hipStream_t s) {
    typedef typename Input<P>::T In;
    const int64_t rows = batches * a.tq;
    const int64_t nqt = (a.tq + BM - 1) / BM;
    if (stage & 1)
        hipLaunchKernelGGL((attn_delta_kernel<In, D>), dim3((unsigned)((rows + 15) / 16)), dim3(256), 0, s, a, rows);
    if (stage & 2) {
        dim3 g1((unsigned)((a.tk + BM - 1) / BM), (unsigned)batches);
        hipLaunchKernelGGL((attn_bwd_dkdv_kernel<P, D>), g1, dim3(256), 0, s, a, (int64_t)0, nqt);
    }
    if (stage & 4) {
        dim3 g2((unsigned)nqt, (unsigned)batches);
        hipLaunchKernelGGL((attn_bwd_dq_kernel<P, D>), g2, dim3(256), 0, s, a);
    }
}

template <int P>
bool dispatch_bwd(const BwdArgs &a, int64_t batches, const ef_attention_desc &desc, int stage, hipStream_t s) {
    switch (desc.features) {
    case 32: launch_bwd<P, 32>(a, batches, stage, s); return true;
    case 64: launch_bwd<P, 64>(a, batches, stage, s); return true;
    default: return false;
    }
}

}  // namespace

ef_status attention_fwd(const ef_attention_desc &d, const void *q, const void *k, const void *v,
                        void *out, float *row_max, float *row_sum, uint64_t *tiles_visited,
                        hipStream_t s) {
    if (d.batches == 0 || d.queries == 0) return EF_OK;
    FwdArgs a;
    a.q = q; a.k = k; a.v = v; a.out = out;
    a.row_max = row_max; a.row_sum = row_sum;
    a.visited = reinterpret_cast<unsigned long long *>(tiles_visited);
    a.mask = make_mask(d);
    a.tq = (int64_t)d.queries;
    a.tk = (int64_t)d.keys;
    a.scale = d.scale;
    a.reverse = d.causal ? 1 : 0;
    const bool fast = d.features == d.values && (d.features == 32 || d.features == 64 || d.features == 128);
    bool ok = false;
    if (fast) {
        switch (d.precision) {
        case EF_PRECISION_FP32: ok = dispatch_fwd<PREC_F32>(a, d.batches, d.features, s); break;
        case EF_PRECISION_BF16X3: ok = dispatch_fwd<PREC_BF16X3>(a, d.batches, d.features, s); break;
        case EF_PRECISION_BF16: ok = dispatch_fwd<PREC_BF16>(a, d.batches, d.features, s); break;
        case EF_PRECISION_FP64: ok = dispatch_fwd<PREC_F64>(a, d.batches, d.features, s); break;
        default: break;
        }
    }
    if (!ok) {
        // Any other shape: the generic kernel (FP64 for double tensors,
        // FP32 for float ones; the tile count is not tracked here).
        const int64_t chunks = ((int64_t)d.values + 7) / 8;
        const size_t threads = d.batches * d.queries * (size_t)chunks;
        if (d.precision == EF_PRECISION_FP64)
            hipLaunchKernelGGL(attn_fwd_generic<double>, dim3(blocks_for(threads, 256)), dim3(256), 0, s,
                               (const double *)q, (const double *)k, (const double *)v, (double *)out,
                               row_max, row_sum, a.mask, (int64_t)d.batches, a.tq, a.tk,
                               (int64_t)d.features, (int64_t)d.values, d.scale);
        else
            hipLaunchKernelGGL(attn_fwd_generic<float>, dim3(blocks_for(threads, 256)), dim3(256), 0, s,
                               (const float *)q, (const float *)k, (const float *)v, (float *)out,
                               row_max, row_sum, a.mask, (int64_t)d.batches, a.tq, a.tk,
                               (int64_t)d.features, (int64_t)d.values, d.scale);
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

size_t attention_bwd_scratch_bytes(const ef_attention_desc &d) {
    return d.batches * d.queries * sizeof(float);
}

ef_status attention_bwd(const ef_attention_desc &d, const void *q, const void *k, const void *v,
                        const void *out, const void *dout, const float *row_max,
                        const float *row_sum, void *dq, void *dk, void *dv, void *scratch,
                        int stages, hipStream_t s) {
    if (d.batches == 0) return EF_OK;
    if (!(d.features == d.values && (d.features == 32 || d.features == 64))) {
        set_error("the attention backward pass takes 32 or 64 features, equal for keys and values");
        return EF_INVALID;
    }
    BwdArgs a;
    a.q = q; a.k = k; a.v = v; a.out = out; a.dout = dout;
    a.row_max = row_max; a.row_sum = row_sum;
    a.delta = static_cast<float *>(scratch);
    a.dq = dq; a.dk = dk; a.dv = dv;
    a.mask = make_mask(d);
    a.tq = (int64_t)d.queries;
    a.tk = (int64_t)d.keys;
    a.scale = d.scale;
    bool ok = false;
    switch (d.precision) {
    case EF_PRECISION_FP32: ok = dispatch_bwd<PREC_F32>(a, d.batches, d, stages, s); break;
    case EF_PRECISION_BF16X3: ok = dispatch_bwd<PREC_BF16X3>(a, d.batches, d, stages, s); break;
    case EF_PRECISION_BF16: ok = dispatch_bwd<PREC_BF16>(a, d.batches, d, stages, s); break;
    case EF_PRECISION_FP64: ok = dispatch_bwd<PREC_F64>(a, d.batches, d, stages, s); break;
    default: break;
    }
    if (!ok) {
        set_error("unknown precision");
        return EF_INVALID;
    }
    EF_HIP(hipGetLastError());
    return EF_OK;
}

}  // namespace ef
