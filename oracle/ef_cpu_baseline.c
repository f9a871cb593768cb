/*
 * TEST INFRASTRUCTURE -- the CPU baseline the bench times beside the GPU:
 * einfold's fused CPU kernel for softmax attention (the fused tile loop of
 * crates/einfold-kernels/src/fused.rs:694-1019 for a q k^T producer,
 * Weight::Exp with a scale and the kept-maximum shift, normalized by the
 * kept groups), restated for speed the way the reference runs it:
 *
 *  - X tiles of 64 query rows are independent tasks spread over threads
 *    (dense::par_tasks, fused.rs:412-428): each task owns its rows' output,
 *    so the bits do not depend on the thread count;
 *  - within a task, key tiles of 64 in order; each 64 x 64 score block and
 *    each 64 x Y weighted sum is one cache-blocked matrix product
 *    (gemm.rs:25-46, which calls matrixmultiply's register-blocked
 *    kernels): here a 4-row x 8-column register block over the block's
 *    inner dimension, FMA where the compiler targets it;
 *  - the online max, rescale and exp per row (fused.rs:857-935), with
 *    libm's exp as the reference's f64::exp.
 *
 * Only the bench's cpu_baseline leg and the tests run it; it is checked
 * against the oracle's exact restatement (ef_oracle.c) within float
 * summation bounds, never used by the product.
 */
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define XT 64
#define JT 64

typedef double v4d __attribute__((vector_size(32)));

/* c[r][0..n) += sum over k < kk of a[r][k] * b[k][0..n) for `rows` rows
 * (a multiple of 4; the callers pad) of a (row stride lda), b rows of
 * stride ldb, c rows of stride ldc; n a multiple of 8.  Register blocks of
 * 4 rows x 8 columns: 8 accumulators, 2 row vectors of b and 4 broadcasts
 * of a per step of k. */
static void block_product(int rows, int kk, int n, const double *a, size_t lda, const double *b, size_t ldb,
                          double *c, size_t ldc) {
    for (int r0 = 0; r0 < rows; r0 += 4) {
        const double *a0 = a + (size_t)r0 * lda;
        for (int c0 = 0; c0 < n; c0 += 8) {
            double *cr = c + (size_t)r0 * ldc + c0;
            v4d acc[4][2];
#pragma GCC unroll 4
            for (int i = 0; i < 4; i++) {
                memcpy(&acc[i][0], cr + (size_t)i * ldc, sizeof(v4d));
                memcpy(&acc[i][1], cr + (size_t)i * ldc + 4, sizeof(v4d));
            }
            const double *bk = b + c0;
            for (int k = 0; k < kk; k++, bk += ldb) {
                v4d b0, b1;
                memcpy(&b0, bk, sizeof b0);
                memcpy(&b1, bk + 4, sizeof b1);
#pragma GCC unroll 4
                for (int i = 0; i < 4; i++) {
                    const double x = a0[(size_t)i * lda + k];
                    const v4d av = {x, x, x, x};
                    acc[i][0] += av * b0;
                    acc[i][1] += av * b1;
                }
            }
#pragma GCC unroll 4
            for (int i = 0; i < 4; i++) {
                memcpy(cr + (size_t)i * ldc, &acc[i][0], sizeof(v4d));
                memcpy(cr + (size_t)i * ldc + 4, &acc[i][1], sizeof(v4d));
            }
        }
    }
}

typedef struct {
    const double *q, *kt, *v;   /* q [nx][ne], kt [ne][nj], v [nj][ny] */
    size_t nx, ne, nj, ny, ny16;
    double scale;
    const size_t *lo, *hi;       /* each query's allowed keys [lo, hi) */
    double *out;                 /* [nx][ny] */
    atomic_size_t next;          /* the next X tile */
    size_t tiles;
} Job;

static void run_tile(Job *jb, size_t t, double *s, double *q_tile, double *kt_tile, double *v_tile, double *acc) {
    const size_t x0 = t * XT, rows = jb->nx - x0 < XT ? jb->nx - x0 : XT;
    const size_t ne = jb->ne, nj = jb->nj, ny = jb->ny, ny16 = jb->ny16;
    double m[XT], z[XT];
    for (size_t r = 0; r < rows; r++) {
        m[r] = -INFINITY;
        z[r] = 0.0;
    }
    memset(acc, 0, sizeof(double) * XT * ny16);
    /* The tile's query rows, zero past nx (whole blocks of 4 rows). */
    memcpy(q_tile, jb->q + x0 * ne, rows * ne * sizeof(double));
    memset(q_tile + rows * ne, 0, (XT - rows) * ne * sizeof(double));
    size_t first = nj, last = 0;
    for (size_t r = 0; r < rows; r++)
        if (jb->lo[x0 + r] < jb->hi[x0 + r]) {
            if (jb->lo[x0 + r] < first) first = jb->lo[x0 + r];
            if (jb->hi[x0 + r] > last) last = jb->hi[x0 + r];
        }
    for (size_t j0 = first - first % JT; j0 < last; j0 += JT) {
        const size_t width = nj - j0 < JT ? nj - j0 : JT;
        /* The tile's keys (zero past nj) and values, packed. */
        for (size_t e = 0; e < ne; e++) {
            memcpy(kt_tile + e * JT, jb->kt + e * nj + j0, width * sizeof(double));
            for (size_t c = width; c < JT; c++) kt_tile[e * JT + c] = 0.0;
        }
        for (size_t c = 0; c < JT; c++)
            for (size_t y = 0; y < ny16; y++) v_tile[c * ny16 + y] = c < width && y < ny ? jb->v[(j0 + c) * ny + y] : 0.0;
        memset(s, 0, sizeof(double) * XT * JT);
        block_product(XT, (int)ne, JT, q_tile, ne, kt_tile, JT, s, JT);
        for (size_t r = rows; r < XT; r++) memset(s + r * JT, 0, JT * sizeof(double));
        for (size_t r = 0; r < rows; r++) {
            const size_t lo = jb->lo[x0 + r] > j0 ? jb->lo[x0 + r] - j0 : 0;
            size_t hi = jb->hi[x0 + r] > j0 ? jb->hi[x0 + r] - j0 : 0;
            if (hi > width) hi = width;
            double *row = s + r * JT;
            double mx = m[r];
            for (size_t c = 0; c < JT; c++) {
                row[c] = (c >= lo && c < hi) ? row[c] * jb->scale : -INFINITY;
                if (row[c] > mx) mx = row[c];
            }
            if (mx == -INFINITY) {
                for (size_t c = 0; c < JT; c++) row[c] = 0.0;
                continue;
            }
            const double rescale = exp(m[r] - mx);
            if (rescale != 1.0) {
                z[r] *= rescale;
                for (size_t y = 0; y < ny16; y++) acc[r * ny16 + y] *= rescale;
            }
            double sum = 0.0;
            for (size_t c = 0; c < JT; c++) {
                row[c] = c >= lo && c < hi ? exp(row[c] - mx) : 0.0;
                sum += row[c];
            }
            z[r] += sum;
            m[r] = mx;
        }
        block_product(XT, JT, (int)ny16, s, JT, v_tile, ny16, acc, ny16);
    }
    for (size_t r = 0; r < rows; r++)
        for (size_t y = 0; y < ny; y++) jb->out[(x0 + r) * ny + y] = acc[r * ny16 + y] / z[r];
}

static void *worker(void *arg) {
    Job *jb = (Job *)arg;
    double *s = aligned_alloc(64, sizeof(double) * XT * JT);
    double *q_tile = aligned_alloc(64, sizeof(double) * XT * (jb->ne ? jb->ne : 1));
    double *kt_tile = aligned_alloc(64, sizeof(double) * (jb->ne ? jb->ne : 1) * JT);
    double *v_tile = aligned_alloc(64, sizeof(double) * JT * jb->ny16);
    double *acc = aligned_alloc(64, sizeof(double) * XT * jb->ny16);
    for (;;) {
        const size_t t = atomic_fetch_add(&jb->next, 1);
        if (t >= jb->tiles) break;
        run_tile(jb, t, s, q_tile, kt_tile, v_tile, acc);
    }
    free(s);
    free(q_tile);
    free(kt_tile);
    free(v_tile);
    free(acc);
    return NULL;
}

/* Softmax attention of one head: out[x] = sum over allowed j of
 * exp(scale q[x].k[j] - max) v[j] / sum of the weights; threads >= 1. */
void ef_cpu_attention_fwd(const double *q, const double *kt, const double *v, size_t nx, size_t ne, size_t nj,
                          size_t ny, double scale, const size_t *lo, const size_t *hi, double *out, int threads) {
    Job jb;
    jb.q = q;
    jb.kt = kt;
    jb.v = v;
    jb.nx = nx;
    jb.ne = ne;
    jb.nj = nj;
    jb.ny = ny;
    jb.ny16 = (ny + 15) / 16 * 16;
    jb.scale = scale;
    jb.lo = lo;
    jb.hi = hi;
    jb.out = out;
    atomic_init(&jb.next, 0);
    jb.tiles = (nx + XT - 1) / XT;
    if (threads < 1) threads = 1;
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
    for (int i = 1; i < threads; i++) pthread_create(&th[i], NULL, worker, &jb);
    worker(&jb);
    for (int i = 1; i < threads; i++) pthread_join(th[i], NULL);
    free(th);
}
