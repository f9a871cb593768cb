/*
 * TEST INFRASTRUCTURE — the CPU oracle for einfold_amd's GPU kernels.
 *
 * A plain-C restatement of the CPU kernels of einfold's EinFoldExec path
 * (crates/einfold-kernels in xqlsystems/einfold).  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg may load this library, and only as
 * the checker: the product path never links or calls it.
 *
 * Every function cites the reference file:line whose algorithm it follows.
 * Arithmetic follows Rust's: `a * b + c` is two roundings (rustc never
 * contracts to FMA), so this file must be compiled with -ffp-contract=off.
 *
 * Parity pins (what checks this oracle, in tests/test_oracle_*.py):
 *   splitmix64          -- known answers of rng.rs:65-78
 *   dense_accumulate    -- numpy float64 einsum within 2(n-1)u*sum|terms|,
 *                          exact metamorphic scale/negate, brute force loops
 *   ranks_of_words      -- numpy unique(return_index, return_inverse)
 *   gustavson           -- brute-force join + group-by on small inputs
 *   fused_tiles         -- numpy float64 masked softmax, tile-count known
 *                          answers of pipeline_vs_sql.rs:734-760, Bounds
 *                          known cases of fused.rs:1116-1177
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#if defined(__FP_FAST_FMA) && !defined(EF_ORACLE_ALLOW_FMA)
/* Contraction is switched off by the build recipe; this is only a guard. */
#endif

/* ------------------------------------------------------------------ */
/* SplitMix64: crates/einfold-testkit/src/rng.rs:17-43                  */
/* ------------------------------------------------------------------ */
uint64_t ef_oracle_seeded(uint64_t seed, uint64_t salt) {
    return (seed * 0x2545f4914f6cdd1dULL) ^ salt;
}

uint64_t ef_oracle_next(uint64_t *state) {
    *state += 0x9e3779b97f4a7c15ULL;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* rng.rs:45-48: 53 random bits in [0, 1). */
double ef_oracle_unit(uint64_t *state) {
    return (double)(ef_oracle_next(state) >> 11) / 9007199254740992.0;
}

void ef_oracle_fill_next(uint64_t seed, uint64_t salt, uint64_t *out, size_t n) {
    uint64_t s = ef_oracle_seeded(seed, salt);
    for (size_t i = 0; i < n; i++) out[i] = ef_oracle_next(&s);
}

/* ------------------------------------------------------------------ */
/* Dense contraction: crates/einfold-kernels/src/dense.rs:706-770       */
/* ------------------------------------------------------------------ */
#define DENSE_XB 256
#define DENSE_YB 2048
#define DENSE_ZB 4096
#define DENSE_FEW_CELLS 4096
#define HELD_VALUED 2

/* out[x][y] = sum_z a[x,z]*b[w][z][y] over VALUED cells of `held`, with
 * w = x / per_batch.  With nx*ny >= FEW_CELLS every cell is one block's,
 * summed from +0.0 in z order (dense.rs:726-751); otherwise z is split in
 * ZB blocks whose partials, each from +0.0, are added to +0.0 in block
 * order (dense.rs:752-768). */
void ef_oracle_dense_accumulate(const double *a, const uint8_t *held, const double *b,
                                size_t nx, size_t nz, size_t ny, size_t per_batch,
                                double *out) {
    if (nx * ny >= DENSE_FEW_CELLS) {
        for (size_t x = 0; x < nx; x++) {
            double *row = out + x * ny;
            for (size_t y = 0; y < ny; y++) row[y] = 0.0;
            size_t w = x / per_batch;
            for (size_t z = 0; z < nz; z++) {
                if (held[x * nz + z] != HELD_VALUED) continue;
                double v = a[x * nz + z];
                const double *brow = b + (w * nz + z) * ny;
                for (size_t y = 0; y < ny; y++) row[y] += v * brow[y];
            }
        }
        return;
    }
    size_t cells = nx * ny;
    for (size_t c = 0; c < cells; c++) out[c] = 0.0;
    double *part = (double *)malloc((cells ? cells : 1) * sizeof(double));
    for (size_t z0 = 0; z0 < nz; z0 += DENSE_ZB) {
        size_t z1 = z0 + DENSE_ZB < nz ? z0 + DENSE_ZB : nz;
        for (size_t c = 0; c < cells; c++) part[c] = 0.0;
        for (size_t z = z0; z < z1; z++) {
            for (size_t x = 0; x < nx; x++) {
                if (held[x * nz + z] != HELD_VALUED) continue;
                double v = a[x * nz + z];
                const double *brow = b + ((x / per_batch) * nz + z) * ny;
                for (size_t y = 0; y < ny; y++) part[x * ny + y] += v * brow[y];
            }
        }
        for (size_t c = 0; c < cells; c++) out[c] += part[c];
    }
    free(part);
}

/* ------------------------------------------------------------------ */
/* Radix ranking: crates/einfold-kernels/src/discriminate.rs:103-156,   */
/* 280-292.  Stable LSD byte sort of (word, row), skipping bytes no     */
/* word varies in; then ranks in word order and each rank's first row. */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t w; uint32_t r; } pair_t;

static void radix_pairs(const uint64_t *words, size_t n, pair_t *pairs, pair_t *next) {
    uint64_t any = 0, all = ~0ULL;
    for (size_t i = 0; i < n; i++) {
        pairs[i].w = words[i];
        pairs[i].r = (uint32_t)i;
        any |= words[i];
        all &= words[i];
    }
    for (int shift = 0; shift < 64; shift += 8) {
        if ((((any ^ all) >> shift) & 0xff) == 0) continue;
        size_t start[257];
        memset(start, 0, sizeof start);
        for (size_t i = 0; i < n; i++) start[((pairs[i].w >> shift) & 0xff) + 1]++;
        for (int d = 0; d < 256; d++) start[d + 1] += start[d];
        for (size_t i = 0; i < n; i++) next[start[(pairs[i].w >> shift) & 0xff]++] = pairs[i];
        pair_t *t = pairs; pairs = next; next = t;
        memcpy(next, pairs, n * sizeof(pair_t));
    }
}

/* rank[i] for every word, first[k] the first row (in sorted, stable
 * order, so the lowest row) holding the k-th distinct word; returns the
 * number of ranks. */
size_t ef_oracle_ranks_of_words(const uint64_t *words, size_t n, uint32_t *rank, uint32_t *first) {
    if (n == 0) return 0;
    pair_t *pairs = (pair_t *)malloc(n * sizeof(pair_t));
    pair_t *next = (pair_t *)malloc(n * sizeof(pair_t));
    radix_pairs(words, n, pairs, next);
    size_t k = 0;
    for (size_t i = 0; i < n; i++) {
        if (i == 0 || pairs[i].w != pairs[i - 1].w) first[k++] = pairs[i].r;
        rank[pairs[i].r] = (uint32_t)(k - 1);
    }
    free(pairs);
    free(next);
    return k;
}

/* ------------------------------------------------------------------ */
/* Counting sorts, tie-breaks and Gustavson's product:                  */
/* crates/einfold-kernels/src/discriminate.rs:295-330, 493-584          */
/* ------------------------------------------------------------------ */

/* discriminate.rs:295-310: `rows` stably by key (< k). */
static void counting(const uint32_t *rows, size_t n, const uint32_t *key, size_t k, uint32_t *out) {
    size_t *start = (size_t *)calloc(k + 1, sizeof(size_t));
    for (size_t i = 0; i < n; i++) start[key[rows[i]] + 1]++;
    for (size_t i = 0; i < k; i++) start[i + 1] += start[i];
    for (size_t i = 0; i < n; i++) out[start[key[rows[i]]]++] = rows[i];
    free(start);
}

/* The sort key of break_ties (discriminate.rs:315-318): NULL first, then
 * the factor's bits as an unsigned integer (Option<u64> ordering). */
static const uint8_t *tie_valid;
static const double *tie_value;
static int tie_cmp_rows(uint32_t p, uint32_t q) {
    int vp = tie_valid[p] != 0, vq = tie_valid[q] != 0;
    if (vp != vq) return vp - vq;
    if (!vp) return 0;
    uint64_t bp, bq;
    memcpy(&bp, &tie_value[p], 8);
    memcpy(&bq, &tie_value[q], 8);
    return bp < bq ? -1 : bp > bq ? 1 : 0;
}

/* Stable insertion-merge sort of one run by its factor bits (Rust's
 * sort_by_key is stable). */
static void stable_sort_run(uint32_t *run, size_t n, uint32_t *tmp) {
    if (n < 2) return;
    size_t h = n / 2;
    stable_sort_run(run, h, tmp);
    stable_sort_run(run + h, n - h, tmp);
    size_t i = 0, j = h, o = 0;
    while (i < h && j < n) tmp[o++] = tie_cmp_rows(run[j], run[i]) < 0 ? run[j++] : run[i++];
    while (i < h) tmp[o++] = run[i++];
    while (j < n) tmp[o++] = run[j++];
    memcpy(run, tmp, n * sizeof(uint32_t));
}

static void break_ties(uint32_t *rows, size_t n, const uint32_t *k1, const uint32_t *k2,
                       const uint8_t *valid, const double *value) {
    uint32_t *tmp = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    tie_valid = valid;
    tie_value = value;
    size_t s = 0;
    while (s < n) {
        size_t e = s + 1;
        while (e < n && k1[rows[s]] == k1[rows[e]] && k2[rows[s]] == k2[rows[e]]) e++;
        if (e - s > 1) stable_sort_run(rows + s, e - s, tmp);
        s = e;
    }
    free(tmp);
}

/*
 * Gustavson's product over ranked rows.  a's rows have x and z ranks, b's
 * z and y ranks; `*_join` says which rows can join (no NULL in a key the
 * join compares with `=`); factors are (valid, value).  Writes the groups
 * in (x, y) order: their ranks, sums and whether any product was valued.
 * Returns the number of groups, or SIZE_MAX if more than `capacity`.
 */
size_t ef_oracle_gustavson(size_t na, const uint32_t *a_x, const uint32_t *a_z, const uint8_t *a_join,
                           const uint8_t *a_valid, const double *a_value,
                           size_t nb, const uint32_t *b_z, const uint32_t *b_y, const uint8_t *b_join,
                           const uint8_t *b_valid, const double *b_value,
                           size_t nx, size_t nz, size_t ny, size_t capacity,
                           uint32_t *g_x, uint32_t *g_y, double *g_sum, uint8_t *g_valued) {
    /* a's joining rows in (x, z) order, duplicates by their factors
     * (discriminate.rs:497-501). */
    uint32_t *rows = (uint32_t *)malloc((na + nb + 1) * sizeof(uint32_t));
    uint32_t *tmp = (uint32_t *)malloc((na + nb + 1) * sizeof(uint32_t));
    size_t n_a = 0;
    for (size_t r = 0; r < na; r++) if (a_join[r]) rows[n_a++] = (uint32_t)r;
    counting(rows, n_a, a_z, nz, tmp);
    uint32_t *a_rows = (uint32_t *)malloc((n_a + 1) * sizeof(uint32_t));
    counting(tmp, n_a, a_x, nx, a_rows);
    break_ties(a_rows, n_a, a_x, a_z, a_valid, a_value);
    /* b's in (z, y) order (discriminate.rs:502-506). */
    size_t n_b = 0;
    for (size_t r = 0; r < nb; r++) if (b_join[r]) rows[n_b++] = (uint32_t)r;
    counting(rows, n_b, b_y, ny, tmp);
    uint32_t *b_rows = (uint32_t *)malloc((n_b + 1) * sizeof(uint32_t));
    counting(tmp, n_b, b_z, nz, b_rows);
    break_ties(b_rows, n_b, b_z, b_y, b_valid, b_value);
    free(rows);
    free(tmp);
    /* Starts of each x's rows of a and each z's rows of b
     * (discriminate.rs:508-519). */
    size_t *a_start = (size_t *)calloc(nx + 1, sizeof(size_t));
    size_t *b_start = (size_t *)calloc(nz + 1, sizeof(size_t));
    for (size_t i = 0; i < n_a; i++) a_start[a_x[a_rows[i]] + 1]++;
    for (size_t i = 0; i < nx; i++) a_start[i + 1] += a_start[i];
    for (size_t i = 0; i < n_b; i++) b_start[b_z[b_rows[i]] + 1]++;
    for (size_t i = 0; i < nz; i++) b_start[i + 1] += b_start[i];

    double *acc = (double *)calloc(ny + 1, sizeof(double));
    uint8_t *valued = (uint8_t *)calloc(ny + 1, 1);
    uint8_t *reached = (uint8_t *)calloc(ny + 1, 1);
    uint32_t *touched = (uint32_t *)malloc((ny + 1) * sizeof(uint32_t));
    size_t groups = 0;
    int overflow = 0;
    /* discriminate.rs:548-583: each x in order; each of its rows of a;
     * each row of b with its z; acc[y] += p * q when both are valued. */
    for (size_t xr = 0; xr < nx && !overflow; xr++) {
        size_t nt = 0;
        for (size_t i = a_start[xr]; i < a_start[xr + 1]; i++) {
            uint32_t r = a_rows[i];
            size_t zr = a_z[r];
            int av = a_valid[r] != 0;
            double p = a_value[r];
            for (size_t j = b_start[zr]; j < b_start[zr + 1]; j++) {
                uint32_t br = b_rows[j];
                size_t yr = b_y[br];
                if (!reached[yr]) { reached[yr] = 1; touched[nt++] = (uint32_t)yr; }
                if (av && b_valid[br]) {
                    acc[yr] += p * b_value[br];
                    valued[yr] = 1;
                }
            }
        }
        /* touched.sort_unstable(): distinct ranks, so any sort is the same. */
        for (size_t i = 1; i < nt; i++) {
            uint32_t t = touched[i];
            size_t k = i;
            while (k > 0 && touched[k - 1] > t) { touched[k] = touched[k - 1]; k--; }
            touched[k] = t;
        }
        for (size_t i = 0; i < nt; i++) {
            uint32_t yr = touched[i];
            if (groups >= capacity) { overflow = 1; break; }
            g_x[groups] = (uint32_t)xr;
            g_y[groups] = yr;
            g_sum[groups] = acc[yr];
            g_valued[groups] = valued[yr];
            groups++;
            acc[yr] = 0.0;
            valued[yr] = 0;
            reached[yr] = 0;
        }
    }
    free(acc); free(valued); free(reached); free(touched);
    free(a_start); free(b_start); free(a_rows); free(b_rows);
    return overflow ? SIZE_MAX : groups;
}

/* ------------------------------------------------------------------ */
/* Row expressions: crates/einfold-ir/src/expression.rs:59-100,         */
/* evaluated per pair as crates/einfold-kernels/src/expression.rs does  */
/* (same element-wise operators, so the same bits).                     */
/* ------------------------------------------------------------------ */
enum {
    EXPR_COLUMN = 0, EXPR_LITERAL = 1,
    EXPR_NEG = 2, EXPR_ABS = 3, EXPR_EXP = 4, EXPR_LN = 5,
    EXPR_ADD = 6, EXPR_SUB = 7, EXPR_MUL = 8, EXPR_DIV = 9, EXPR_POW = 10
};

typedef struct {
    uint32_t code;
    uint32_t operand;   /* EXPR_COLUMN: 0 = q (row side), 1 = k (column side) */
    uint32_t column;    /* EXPR_COLUMN: slot of the placed column */
    uint32_t pad;
    double value;       /* EXPR_LITERAL */
} ef_oracle_expr;

/* Evaluates a postfix expression; `cols[o][slot]` are the column values of
 * operand o's row of this pair. */
static double eval_expr(const ef_oracle_expr *e, size_t n, const double *row_vals,
                        const double *col_vals) {
    double stack[64];
    int sp = 0;
    for (size_t i = 0; i < n; i++) {
        const ef_oracle_expr *op = &e[i];
        switch (op->code) {
        case EXPR_COLUMN:
            stack[sp++] = op->operand == 0 ? row_vals[op->column] : col_vals[op->column];
            break;
        case EXPR_LITERAL: stack[sp++] = op->value; break;
        case EXPR_NEG: stack[sp - 1] = -stack[sp - 1]; break;
        case EXPR_ABS: stack[sp - 1] = fabs(stack[sp - 1]); break;
        case EXPR_EXP: stack[sp - 1] = exp(stack[sp - 1]); break;
        case EXPR_LN: stack[sp - 1] = log(stack[sp - 1]); break;
        default: {
            double b = stack[--sp], a = stack[sp - 1], r;
            switch (op->code) {
            case EXPR_ADD: r = a + b; break;
            case EXPR_SUB: r = a - b; break;
            case EXPR_MUL: r = a * b; break;
            case EXPR_DIV: r = a / b; break;
            default: r = pow(a, b); break;
            }
            stack[sp - 1] = r;
        }
        }
    }
    return stack[0];
}

/* Evaluates an expression over an array of single values (the Weight::Of
 * path, where every column read is the score): for tests of the evaluator. */
void ef_oracle_eval(const ef_oracle_expr *e, size_t n_ops, const double *x, size_t n, double *out) {
    for (size_t i = 0; i < n; i++) {
        double v[8] = {x[i], x[i], x[i], x[i], x[i], x[i], x[i], x[i]};
        out[i] = eval_expr(e, n_ops, v, v);
    }
}

/* ------------------------------------------------------------------ */
/* The fused pipeline's tiles: crates/einfold-kernels/src/fused.rs      */
/* 694-1019 (tile), 1024-1086 (overflows, Bounds::hold).                */
/* ------------------------------------------------------------------ */
#define XT 64
#define JT 64

enum { SCALE_ONE = 0, SCALE_TIMES = 1, SCALE_OVER = 2 };
enum { NORM_KEPT_GROUPS = 0, NORM_JOINED_ROWS = 1, NORM_NONE = 2 };

typedef struct {
    /* Weight: exp (Weight::Exp) when weight_expr is NULL, else Weight::Of. */
    uint32_t scale_kind;
    uint32_t shifted;        /* Shift::KeptMaximum */
    uint32_t normalizer;
    uint32_t n_weight_ops;
    double scale;
    const ef_oracle_expr *weight_expr;
    /* Producer: a product q[x][e]*kt[e][j] when score_expr is NULL, else an
     * expression of the placed columns, q-side columns `q_cols` matrices
     * [x][e] and k-side `k_cols` matrices [e][j]. */
    const ef_oracle_expr *score_expr;
    uint32_t n_score_ops;
    uint32_t n_q_cols;
    uint32_t n_k_cols;
    uint32_t pad;
    const double *const *q_cols;
    const double *const *k_cols;
} ef_oracle_pipeline;

static double scale_apply(const ef_oracle_pipeline *p, double s) {
    switch (p->scale_kind) {
    case SCALE_TIMES: return s * p->scale;
    case SCALE_OVER: return s / p->scale;
    default: return s;
    }
}

/* fused.rs:1024-1028 */
static int overflows(double z, const double *a, int special_weight, const uint8_t *special_value, size_t ny) {
    if (special_weight) return 0;
    if (!isfinite(z)) return 1;
    for (size_t y = 0; y < ny; y++)
        if (!special_value[y] && !isfinite(a[y])) return 1;
    return 0;
}

/* fused.rs:1058-1086: Bounds::hold.  has_z = 0 for no normalizer. */
int ef_oracle_bounds_hold(double heaviest, double lightest, double largest, double smallest,
                          size_t terms_n, int has_z, double z, const double *results,
                          const uint8_t *special, size_t n) {
    double terms = 2.0 * (double)terms_n;
    int finite = isfinite(heaviest * largest * terms);
    if (finite && has_z) {
        double quotient = heaviest / fabs(z);
        finite = isfinite(quotient) && isfinite(quotient * largest * terms);
    }
    double inverse = has_z ? 1.0 / fabs(z) : 1.0;
    double tiny;
    uint64_t one = 1;
    memcpy(&tiny, &one, 8);
    double lost = terms * tiny * (inverse + largest + 1.0);
    double rounding = DBL_EPSILON / 4.0;
    double normal = 2.0 * DBL_MIN;
    int no_underflow = smallest == INFINITY;
    if (!no_underflow) {
        no_underflow = lightest * smallest >= normal;
        if (no_underflow && has_z) {
            double quotient = lightest / fabs(z);
            no_underflow = quotient >= normal && quotient * smallest >= normal;
        }
    }
    if (!finite) return 0;
    if (no_underflow) return 1;
    for (size_t i = 0; i < n; i++)
        if (!special[i] && !(lost <= rounding * fabs(results[i]))) return 0;
    return 1;
}

/*
 * All X tiles of a pipeline, one after another (each tile's bits are its
 * own, so the order of tiles never shows).  Shapes: nx X positions, ne E
 * positions, nj J positions, ny Y positions; q [nx][ne], kt [ne][nj],
 * v [nj][ny]; `lo`/`hi` each X position's allowed J interval (global J
 * positions).  Writes out[nx][ny] (rows with an empty interval are left as
 * the tile computed them: the caller drops them), the tiles visited, and
 * returns 1 when some tile overflowed (the pipeline declines).
 */
int ef_oracle_fused(const ef_oracle_pipeline *p, size_t nx, size_t ne, size_t nj, size_t ny,
                    const double *q, const double *kt, const double *v,
                    const size_t *lo_x, const size_t *hi_x, double *out, size_t *visited_out) {
    int overflowed = 0;
    size_t visited = 0;
    int unnormalized = p->normalizer == NORM_NONE;
    int shifted = p->weight_expr == NULL && p->shifted;
    int mapped = p->weight_expr != NULL;
    /* Per J position: finite, largest finite magnitude, smallest nonzero
     * finite magnitude (fused.rs:371-384). */
    uint8_t *finite = (uint8_t *)malloc(nj + 1);
    double *largest = (double *)malloc((nj + 1) * sizeof(double));
    double *smallest = (double *)malloc((nj + 1) * sizeof(double));
    for (size_t j = 0; j < nj; j++) {
        int f = 1;
        double l = 0.0, s = INFINITY;
        for (size_t y = 0; y < ny; y++) {
            double x = v[j * ny + y];
            if (!isfinite(x)) { f = 0; continue; }
            double m = fabs(x);
            l = fmax(l, m);
            if (m != 0.0) s = fmin(s, m);
        }
        finite[j] = (uint8_t)f;
        largest[j] = l;
        smallest[j] = s;
    }
    double *block = (double *)malloc(XT * JT * sizeof(double));
    double *m = (double *)malloc(XT * sizeof(double));
    double *z = (double *)malloc(XT * sizeof(double));
    double *a = (double *)malloc(XT * (ny ? ny : 1) * sizeof(double));
    uint8_t *special_weight = (uint8_t *)malloc(XT);
    uint8_t *special_value = (uint8_t *)malloc(XT * (ny ? ny : 1));
    double *high = (double *)malloc(XT * sizeof(double));
    double *light = (double *)malloc(XT * sizeof(double));
    size_t *terms = (size_t *)malloc(XT * sizeof(size_t));
    double *row_largest = (double *)malloc(XT * sizeof(double));
    double *row_smallest = (double *)malloc(XT * sizeof(double));
    double *low = (double *)malloc(XT * sizeof(double));
    size_t span_lo[XT], span_hi[XT];
    double qv[16], kv[16];

    for (size_t x0 = 0; x0 < nx; x0 += XT) {
        size_t x1 = x0 + XT < nx ? x0 + XT : nx;
        size_t rows = x1 - x0;
        for (size_t r = 0; r < rows; r++) {
            m[r] = shifted ? -INFINITY : 0.0;
            z[r] = 0.0;
            special_weight[r] = 0;
            high[r] = -INFINITY;
            light[r] = INFINITY;
            terms[r] = 0;
            row_largest[r] = 0.0;
            row_smallest[r] = INFINITY;
            low[r] = INFINITY;
            for (size_t y = 0; y < ny; y++) { a[r * ny + y] = 0.0; special_value[r * ny + y] = 0; }
        }
        /* fused.rs:750-755 */
        size_t first = nj, last = 0;
        for (size_t x = x0; x < x1; x++) {
            if (lo_x[x] < hi_x[x]) {
                if (lo_x[x] < first) first = lo_x[x];
                if (hi_x[x] > last) last = hi_x[x];
            }
        }
        for (size_t j0 = first - first % JT; j0 < last; j0 += JT) {
            size_t width = JT < nj - j0 ? JT : nj - j0;
            int any = 0, full = 1;
            for (size_t r = 0; r < rows; r++) {
                size_t l = lo_x[x0 + r], h = hi_x[x0 + r];
                l = l < j0 ? j0 : (l > j0 + width ? j0 + width : l);
                h = h < j0 ? j0 : (h > j0 + width ? j0 + width : h);
                span_lo[r] = l - j0;
                span_hi[r] = h - j0;
                if (span_lo[r] < span_hi[r]) any = 1;
                if (!(span_lo[r] == 0 && span_hi[r] == width)) full = 0;
            }
            if (!any) continue;
            visited++;
            /* The producer's block (fused.rs:771-811). */
            if (p->score_expr == NULL) {
                for (size_t r = 0; r < rows; r++)
                    for (size_t c = 0; c < width; c++) {
                        double s = 0.0;
                        for (size_t e = 0; e < ne; e++) s += q[(x0 + r) * ne + e] * kt[e * nj + j0 + c];
                        block[r * width + c] = s;
                    }
            } else {
                for (size_t r = 0; r < rows; r++)
                    for (size_t c = 0; c < width; c++) {
                        double s = 0.0;
                        for (size_t e = 0; e < ne; e++) {
                            for (uint32_t i = 0; i < p->n_q_cols; i++) qv[i] = p->q_cols[i][(x0 + r) * ne + e];
                            for (uint32_t i = 0; i < p->n_k_cols; i++) kv[i] = p->k_cols[i][e * nj + j0 + c];
                            double t = eval_expr(p->score_expr, p->n_score_ops, qv, kv);
                            s = e == 0 ? t : s + t;
                        }
                        block[r * width + c] = s;
                    }
            }
            /* Weights and each row's state (fused.rs:812-909). */
            for (size_t r = 0; r < rows; r++) {
                size_t lo = span_lo[r], hi = span_hi[r];
                if (lo >= hi) continue;
                terms[r] += hi - lo;
                for (size_t j = j0 + lo; j < j0 + hi; j++) {
                    row_largest[r] = fmax(row_largest[r], largest[j]);
                    row_smallest[r] = fmin(row_smallest[r], smallest[j]);
                }
                double *sc = block + r * width + lo;
                size_t n = hi - lo;
                if (mapped) {
                    for (size_t i = 0; i < n; i++) {
                        double s1[8] = {sc[i], sc[i], sc[i], sc[i], sc[i], sc[i], sc[i], sc[i]};
                        sc[i] = eval_expr(p->weight_expr, p->n_weight_ops, s1, s1);
                    }
                    for (size_t i = 0; i < n; i++)
                        if (!isfinite(sc[i]) || !finite[j0 + lo + i]) overflowed = 1;
                    for (size_t i = 0; i < n; i++) {
                        double w = fabs(sc[i]);
                        high[r] = fmax(high[r], w);
                        if (w != 0.0) light[r] = fmin(light[r], w);
                    }
                    /* Rust's f64 Sum folds from -0.0. */
                    double sum = -0.0;
                    for (size_t i = 0; i < n; i++) sum += sc[i];
                    z[r] += sum;
                    continue;
                }
                for (size_t i = 0; i < n; i++) {
                    sc[i] = scale_apply(p, sc[i]);
                    if (isfinite(sc[i])) {
                        high[r] = fmax(high[r], sc[i]);
                        light[r] = fmin(light[r], sc[i]);
                    }
                }
                for (size_t i = 0; i < n; i++) {
                    size_t j = j0 + lo + i;
                    if (finite[j]) continue;
                    int inf = 0;
                    for (size_t y = 0; y < ny; y++) if (isinf(v[j * ny + y])) inf = 1;
                    if (inf && isfinite(sc[i])) low[r] = fmin(low[r], sc[i]);
                }
                double shift = 0.0;
                if (shifted) {
                    shift = m[r];
                    for (size_t i = 0; i < n; i++)
                        if (sc[i] > shift || isnan(sc[i])) shift = sc[i];
                }
                if (shift == -INFINITY) {
                    for (size_t i = 0; i < n; i++) sc[i] = 0.0;
                    continue;
                }
                double rescale = exp(m[r] - shift);
                double sum = 0.0;
                for (size_t i = 0; i < n; i++) {
                    sc[i] = exp(sc[i] - shift);
                    sum += sc[i];
                }
                int special_weights = 0;
                for (size_t i = 0; i < n; i++) if (!isfinite(sc[i])) special_weights = 1;
                if (shifted && rescale != 1.0) {
                    if (overflows(z[r], a + r * ny, special_weight[r], special_value + r * ny, ny)) overflowed = 1;
                    z[r] *= rescale;
                    for (size_t y = 0; y < ny; y++) a[r * ny + y] *= rescale;
                }
                special_weight[r] |= (uint8_t)special_weights;
                int special_values = 0;
                for (size_t i = 0; i < n; i++) {
                    size_t j = j0 + lo + i;
                    if (finite[j]) continue;
                    special_values = 1;
                    for (size_t y = 0; y < ny; y++)
                        if (!isfinite(v[j * ny + y])) special_value[r * ny + y] = 1;
                }
                if (unnormalized && (special_weights || special_values)) overflowed = 1;
                z[r] += sum;
                m[r] = shift;
            }
            /* a += w . v over allowed pairs (fused.rs:910-935). */
            for (size_t r = 0; r < rows; r++) {
                size_t lo = full ? 0 : span_lo[r], hi = full ? width : span_hi[r];
                if (hi < lo) hi = lo;
                for (size_t i = lo; i < hi; i++) {
                    double w = block[r * width + i];
                    const double *vr = v + (j0 + i) * ny;
                    for (size_t y = 0; y < ny; y++) a[r * ny + y] += w * vr[y];
                }
            }
        }
        /* Normalize (fused.rs:941-964). */
        if (p->normalizer == NORM_NONE) {
            for (size_t r = 0; r < rows; r++)
                if (shifted && m[r] == -INFINITY)
                    for (size_t y = 0; y < ny; y++) a[r * ny + y] = NAN;
        } else {
            for (size_t r = 0; r < rows; r++) {
                double *row = a + r * ny;
                if (overflows(z[r], row, special_weight[r], special_value + r * ny, ny)) overflowed = 1;
                for (size_t y = 0; y < ny; y++) row[y] /= z[r];
                if (z[r] != 0.0 && overflows(z[r], row, special_weight[r], special_value + r * ny, ny)) overflowed = 1;
            }
        }
        /* Underflow and overflow guards (fused.rs:980-1013). */
        for (size_t r = 0; r < rows; r++) {
            double heaviest;
            if (mapped) heaviest = high[r];
            else if (shifted) heaviest = exp(high[r] - m[r]);
            else heaviest = exp(high[r]);
            if (special_weight[r] || high[r] == -INFINITY || heaviest == 0.0) continue;
            double lightest;
            if (mapped) lightest = light[r];
            else if (shifted) lightest = exp(light[r] - m[r]);
            else lightest = exp(light[r]);
            if (!ef_oracle_bounds_hold(heaviest, lightest, row_largest[r], row_smallest[r], terms[r],
                                       !unnormalized, z[r], a + r * ny, special_value + r * ny, ny))
                overflowed = 1;
            if (low[r] != INFINITY) {
                double lw = shifted ? exp(low[r] - m[r]) : exp(low[r]);
                double normal = 2.0 * DBL_MIN;
                int ok = lw >= normal && (unnormalized || lw / fabs(z[r]) >= normal);
                if (!ok) overflowed = 1;
            }
        }
        for (size_t r = 0; r < rows; r++)
            for (size_t y = 0; y < ny; y++) out[(x0 + r) * ny + y] = a[r * ny + y];
    }
    free(finite); free(largest); free(smallest); free(block); free(m); free(z); free(a);
    free(special_weight); free(special_value); free(high); free(light); free(terms);
    free(row_largest); free(row_smallest); free(low);
    *visited_out = visited;
    return overflowed;
}
