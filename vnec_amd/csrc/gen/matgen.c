/*
 * matgen.c — seeded synthetic sparse-matrix generators for the VNEC SpMV
 * workloads (R-MAT, 27-point stencil, skewed R-MAT with dense rows, the
 * 512-matrix batch, and the selector training mix).
 *
 * Every generator draws from one xoshiro256** stream seeded through
 * splitmix64, in a fixed order: structure first, then the values in CSR
 * order, then x.  The same seed therefore yields the same bytes on every
 * machine, so the CPU oracle and the GPU path see identical inputs.
 *
 * Output is CSR with int64 row pointers, int32 column indices (sorted and
 * de-duplicated within each row), FP64 values and an FP64 x of length ncols.
 * The buffers are malloc'ed here and released with vg_free().
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int64_t nrows, ncols, nnz;
    int64_t *ptr;   /* nrows + 1 */
    int32_t *col;   /* nnz */
    double *val;    /* nnz */
    double *x;      /* ncols * nrhs */
    int64_t nrhs;
    /* batched: block-diagonal concatenation of nmat matrices */
    int64_t nmat;
    int64_t *row_off; /* nmat + 1 */
    int64_t *col_off; /* nmat + 1 */
} vg_csr;

/* ---------------------------------------------------------------- RNG */
typedef struct { uint64_t s[4]; } vg_rng;

static uint64_t splitmix64(uint64_t *st)
{
    uint64_t z = (*st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(vg_rng *r, uint64_t seed)
{
    uint64_t st = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&st);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(vg_rng *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* uniform double in [0, 1) with 53 random bits */
static inline double rng_u01(vg_rng *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }
/* uniform double in [-1, 1) */
static inline double rng_pm1(vg_rng *r) { return 2.0 * rng_u01(r) - 1.0; }
/* uniform integer in [0, n) (Lemire's multiply-shift; bias < n / 2^64) */
static inline uint64_t rng_below(vg_rng *r, uint64_t n)
{
    return (uint64_t)(((unsigned __int128)rng_next(r) * n) >> 64);
}

/* Poisson(mu) by inversion; mu <= ~700 keeps exp(-mu) representable. */
static int64_t rng_poisson(vg_rng *r, double mu)
{
    double u = rng_u01(r);
    double p = exp(-mu), cdf = p;
    int64_t k = 0;
    while (u > cdf && k < 100000) {
        k++;
        p *= mu / (double)k;
        cdf += p;
        if (p == 0.0 && cdf < u) break;
    }
    return k;
}

/* ---------------------------------------------------- key sort helpers */
/* LSD radix sort of 64-bit keys, 16 bits per pass; tmp has n entries. */
static void radix_sort_u64(uint64_t *a, uint64_t *tmp, int64_t n, int key_bits)
{
    int passes = (key_bits + 15) / 16;
    size_t *cnt = (size_t *)malloc(sizeof(size_t) * 65536);
    for (int p = 0; p < passes; p++) {
        int sh = 16 * p;
        memset(cnt, 0, sizeof(size_t) * 65536);
        for (int64_t i = 0; i < n; i++) cnt[(a[i] >> sh) & 0xFFFF]++;
        size_t run = 0;
        for (int b = 0; b < 65536; b++) { size_t c = cnt[b]; cnt[b] = run; run += c; }
        for (int64_t i = 0; i < n; i++) tmp[cnt[(a[i] >> sh) & 0xFFFF]++] = a[i];
        uint64_t *sw = a; a = tmp; tmp = sw;
    }
    if (passes & 1) memcpy(tmp, a, sizeof(uint64_t) * (size_t)n); /* result back in caller's a */
    free(cnt);
}

static int bits_for(uint64_t n)
{
    int b = 1;
    while (b < 63 && (1ull << b) < n) b++;
    return b;
}

static vg_csr *csr_alloc(int64_t nrows, int64_t ncols, int64_t nnz)
{
    vg_csr *m = (vg_csr *)calloc(1, sizeof(vg_csr));
    m->nrows = nrows;
    m->ncols = ncols;
    m->nnz = nnz;
    m->nrhs = 1;
    m->ptr = (int64_t *)calloc((size_t)nrows + 1, sizeof(int64_t));
    m->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    m->val = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    return m;
}

/* Sorted, de-duplicated (row << cb | col) keys -> CSR structure. */
static vg_csr *csr_from_keys(uint64_t *keys, int64_t n, int64_t nrows, int64_t ncols, int cb)
{
    int64_t u = 0;
    for (int64_t i = 0; i < n; i++)
        if (i == 0 || keys[i] != keys[i - 1]) keys[u++] = keys[i];
    vg_csr *m = csr_alloc(nrows, ncols, u);
    uint64_t cmask = (1ull << cb) - 1;
    for (int64_t i = 0; i < u; i++) {
        m->ptr[(keys[i] >> cb) + 1]++;
        m->col[i] = (int32_t)(keys[i] & cmask);
    }
    for (int64_t r = 0; r < nrows; r++) m->ptr[r + 1] += m->ptr[r];
    return m;
}

static void fill_values_pm1(vg_csr *m, vg_rng *r)
{
    for (int64_t i = 0; i < m->nnz; i++) m->val[i] = rng_pm1(r);
}

static void fill_x_pm1(vg_csr *m, vg_rng *r, int64_t nrhs)
{
    m->nrhs = nrhs;
    m->x = (double *)malloc(sizeof(double) * (size_t)(m->ncols * nrhs > 0 ? m->ncols * nrhs : 1));
    for (int64_t i = 0; i < m->ncols * nrhs; i++) m->x[i] = rng_pm1(r);
}

/* ------------------------------------------------------------- R-MAT */
static void rmat_edges(vg_rng *r, int scale, int64_t nedges, double a, double b, double c,
                       uint64_t *keys, int cb)
{
    double ab = a + b, abc = a + b + c;
    for (int64_t e = 0; e < nedges; e++) {
        uint64_t row = 0, col = 0;
        for (int l = 0; l < scale; l++) {
            double u = rng_u01(r);
            uint64_t bit = 1ull << (scale - 1 - l);
            if (u < a) {
            } else if (u < ab) {
                col |= bit;
            } else if (u < abc) {
                row |= bit;
            } else {
                row |= bit;
                col |= bit;
            }
        }
        keys[e] = (row << cb) | col;
    }
}

/* R-MAT, scale s, edge factor ef; duplicates merged, self-loops kept. */
vg_csr *vg_rmat(int scale, int ef, double a, double b, double c, uint64_t seed)
{
    vg_rng r;
    rng_seed(&r, seed);
    int64_t n = 1ll << scale;
    int64_t ne = (int64_t)ef * n;
    int cb = scale;
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)ne);
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)ne);
    rmat_edges(&r, scale, ne, a, b, c, keys, cb);
    radix_sort_u64(keys, tmp, ne, 2 * cb);
    free(tmp);
    vg_csr *m = csr_from_keys(keys, ne, n, n, cb);
    free(keys);
    fill_values_pm1(m, &r);
    fill_x_pm1(m, &r, 1);
    return m;
}

/* R-MAT scale s plus `ndense` rows holding `dense_len` distinct uniform
 * columns each (rows drawn uniformly, distinct); merged with the R-MAT edges. */
vg_csr *vg_skew(int scale, int ef, double a, double b, double c, int ndense, int64_t dense_len,
                uint64_t seed)
{
    vg_rng r;
    rng_seed(&r, seed);
    int64_t n = 1ll << scale;
    if (dense_len > n) dense_len = n;
    int64_t ne = (int64_t)ef * n;
    int64_t tot = ne + (int64_t)ndense * dense_len;
    int cb = scale;
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)tot);
    rmat_edges(&r, scale, ne, a, b, c, keys, cb);
    /* distinct dense rows */
    int64_t *drows = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ndense > 0 ? ndense : 1));
    for (int d = 0; d < ndense; d++) {
        for (;;) {
            int64_t cand = (int64_t)rng_below(&r, (uint64_t)n);
            int dup = 0;
            for (int e = 0; e < d; e++) dup |= (drows[e] == cand);
            if (!dup) { drows[d] = cand; break; }
        }
    }
    /* Floyd's sampling of dense_len distinct columns per dense row */
    uint64_t *bm = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)((n + 63) / 64));
    int64_t k = ne;
    for (int d = 0; d < ndense; d++) {
        memset(bm, 0, sizeof(uint64_t) * (size_t)((n + 63) / 64));
        for (int64_t j = n - dense_len; j < n; j++) {
            int64_t t = (int64_t)rng_below(&r, (uint64_t)(j + 1));
            int64_t pick = ((bm[t >> 6] >> (t & 63)) & 1) ? j : t;
            bm[pick >> 6] |= 1ull << (pick & 63);
            keys[k++] = ((uint64_t)drows[d] << cb) | (uint64_t)pick;
        }
    }
    free(bm);
    free(drows);
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)tot);
    radix_sort_u64(keys, tmp, tot, 2 * cb);
    free(tmp);
    vg_csr *m = csr_from_keys(keys, tot, n, n, cb);
    free(keys);
    fill_values_pm1(m, &r);
    fill_x_pm1(m, &r, 1);
    return m;
}

/* 27-point stencil on an n^3 grid, row (i*n + j)*n + k, truncated at the
 * boundary.  Diagonal 26 + U[0,1), off-diagonals U[-1,0). */
vg_csr *vg_stencil27(int n, uint64_t seed)
{
    vg_rng r;
    rng_seed(&r, seed);
    int64_t N = (int64_t)n * n * n;
    int64_t per = 3ll * n - 2;
    int64_t nnz = per * per * per;
    vg_csr *m = csr_alloc(N, N, nnz);
    int64_t p = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) {
                int64_t row = ((int64_t)i * n + j) * n + k;
                for (int di = -1; di <= 1; di++) {
                    int ii = i + di;
                    if (ii < 0 || ii >= n) continue;
                    for (int dj = -1; dj <= 1; dj++) {
                        int jj = j + dj;
                        if (jj < 0 || jj >= n) continue;
                        for (int dk = -1; dk <= 1; dk++) {
                            int kk = k + dk;
                            if (kk < 0 || kk >= n) continue;
                            m->col[p++] = (int32_t)(((int64_t)ii * n + jj) * n + kk);
                        }
                    }
                }
                m->ptr[row + 1] = p;
            }
    for (int64_t row = 0; row < N; row++)
        for (int64_t q = m->ptr[row]; q < m->ptr[row + 1]; q++) {
            double u = rng_u01(&r);
            m->val[q] = (m->col[q] == row) ? 26.0 + u : u - 1.0;
        }
    fill_x_pm1(m, &r, 1);
    return m;
}

/* Grid stencil on an nx x ny x nz grid (nz = 1: two-dimensional), row
 * (i*ny + j)*nz + k, truncated at the boundary.  pts = 5 or 9 (2-D star /
 * box), 7 or 27 (3-D star / box).  Diagonal (pts - 1) + U[0,1),
 * off-diagonals U[-1,0). */
vg_csr *vg_stencil(int nx, int ny, int nz, int pts, uint64_t seed)
{
    const int box = pts == 9 || pts == 27;
    const int dz = nz > 1 ? 1 : 0;
    if (nx < 1 || ny < 1 || nz < 1 || !(pts == 5 || pts == 9 || pts == 7 || pts == 27) ||
        ((pts == 5 || pts == 9) && nz != 1) || ((pts == 7 || pts == 27) && nz < 2))
        return NULL;
    vg_rng r;
    rng_seed(&r, seed);
    int64_t N = (int64_t)nx * ny * nz;
    vg_csr *m = csr_alloc(N, N, N * pts);
    int64_t p = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                int64_t row = ((int64_t)i * ny + j) * nz + k;
                for (int di = -1; di <= 1; di++)
                    for (int dj = -1; dj <= 1; dj++)
                        for (int dk = -dz; dk <= dz; dk++) {
                            int nzero = (di != 0) + (dj != 0) + (dk != 0);
                            if (!box && nzero > 1) continue;
                            int ii = i + di, jj = j + dj, kk = k + dk;
                            if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0 || kk >= nz) continue;
                            m->col[p++] = (int32_t)(((int64_t)ii * ny + jj) * nz + kk);
                        }
                m->ptr[row + 1] = p;
            }
    m->nnz = p;
    for (int64_t row = 0; row < N; row++)
        for (int64_t q = m->ptr[row]; q < m->ptr[row + 1]; q++) {
            double u = rng_u01(&r);
            m->val[q] = (m->col[q] == row) ? (double)(pts - 1) + u : u - 1.0;
        }
    fill_x_pm1(m, &r, 1);
    return m;
}

/* --------------------------------------------- per-row random patterns */
static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Build a CSR whose row i gets lens[i] uniform columns in [0, ncols),
 * sorted and de-duplicated. */
static vg_csr *csr_random_rows(vg_rng *r, int64_t nrows, int64_t ncols, const int64_t *lens)
{
    int64_t cap = 0;
    for (int64_t i = 0; i < nrows; i++) cap += lens[i];
    vg_csr *m = csr_alloc(nrows, ncols, cap);
    int64_t p = 0;
    for (int64_t i = 0; i < nrows; i++) {
        int64_t b = p;
        for (int64_t k = 0; k < lens[i]; k++) m->col[p++] = (int32_t)rng_below(r, (uint64_t)ncols);
        int64_t len = p - b;
        if (len > 1) {
            if (len <= 32) {
                for (int64_t a = b + 1; a < p; a++) {
                    int32_t v = m->col[a];
                    int64_t q = a - 1;
                    while (q >= b && m->col[q] > v) { m->col[q + 1] = m->col[q]; q--; }
                    m->col[q + 1] = v;
                }
            } else {
                qsort(m->col + b, (size_t)len, sizeof(int32_t), cmp_i32);
            }
            int64_t u = b + 1;
            for (int64_t a = b + 1; a < p; a++)
                if (m->col[a] != m->col[u - 1]) m->col[u++] = m->col[a];
            p = u;
        }
        m->ptr[i + 1] = p;
    }
    m->nnz = p;
    return m;
}

/* One matrix of the batch: rows log-uniform in [rmin, rmax], square,
 * nnz/row ~ Poisson(mu), mu ~ U[mu_lo, mu_hi], uniform columns. */
static vg_csr *batch_member(uint64_t seed, int64_t rmin, int64_t rmax, double mu_lo, double mu_hi,
                            int64_t nrhs)
{
    vg_rng r;
    rng_seed(&r, seed);
    double l0 = log((double)rmin), l1 = log((double)rmax);
    int64_t n = (int64_t)floor(exp(l0 + (l1 - l0) * rng_u01(&r)));
    if (n < rmin) n = rmin;
    if (n > rmax) n = rmax;
    double mu = mu_lo + (mu_hi - mu_lo) * rng_u01(&r);
    int64_t *lens = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) lens[i] = rng_poisson(&r, mu);
    vg_csr *m = csr_random_rows(&r, n, n, lens);
    free(lens);
    fill_values_pm1(m, &r);
    fill_x_pm1(m, &r, nrhs);
    return m;
}

void vg_free(vg_csr *m)
{
    if (!m) return;
    free(m->ptr);
    free(m->col);
    free(m->val);
    free(m->x);
    free(m->row_off);
    free(m->col_off);
    free(m);
}

/* Batch of nmat independent matrices, concatenated block-diagonally:
 * column indices are stored LOCAL to each member (0..n_m-1); row_off and
 * col_off give each member's first row / first x entry.  Member m is drawn
 * from its own stream seeded with splitmix64(seed + m), so members can be
 * generated in parallel and the result does not depend on thread count.
 * x holds each member's x (n_m * nrhs entries, row-major [col][rhs]). */
vg_csr *vg_batch(int nmat, int64_t rmin, int64_t rmax, double mu_lo, double mu_hi, int64_t nrhs,
                 uint64_t seed)
{
    vg_csr **parts = (vg_csr **)malloc(sizeof(vg_csr *) * (size_t)nmat);
#pragma omp parallel for schedule(dynamic, 1)
    for (int m = 0; m < nmat; m++) {
        uint64_t st = seed + (uint64_t)m;
        parts[m] = batch_member(splitmix64(&st), rmin, rmax, mu_lo, mu_hi, nrhs);
    }
    int64_t R = 0, C = 0, Z = 0;
    for (int m = 0; m < nmat; m++) { R += parts[m]->nrows; C += parts[m]->ncols; Z += parts[m]->nnz; }
    vg_csr *out = csr_alloc(R, C, Z);
    out->nmat = nmat;
    out->nrhs = nrhs;
    out->row_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nmat + 1));
    out->col_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nmat + 1));
    out->x = (double *)malloc(sizeof(double) * (size_t)(C * nrhs > 0 ? C * nrhs : 1));
    int64_t ro = 0, co = 0, zo = 0;
    for (int m = 0; m < nmat; m++) {
        vg_csr *p = parts[m];
        out->row_off[m] = ro;
        out->col_off[m] = co;
        for (int64_t i = 0; i < p->nrows; i++) out->ptr[ro + i + 1] = zo + p->ptr[i + 1];
        memcpy(out->col + zo, p->col, sizeof(int32_t) * (size_t)p->nnz);
        memcpy(out->val + zo, p->val, sizeof(double) * (size_t)p->nnz);
        memcpy(out->x + co * nrhs, p->x, sizeof(double) * (size_t)(p->ncols * nrhs));
        ro += p->nrows;
        co += p->ncols;
        zo += p->nnz;
        vg_free(p);
    }
    out->row_off[nmat] = ro;
    out->col_off[nmat] = co;
    free(parts);
    return out;
}

/* ------------------------------------------- selector training patterns */
/* Banded: row i holds every column in [i - hb, i + hb] within range. */
vg_csr *vg_banded(int64_t n, int hb, uint64_t seed)
{
    vg_rng r;
    rng_seed(&r, seed);
    int64_t cap = n * (2ll * hb + 1);
    vg_csr *m = csr_alloc(n, n, cap);
    int64_t p = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t lo = i - hb < 0 ? 0 : i - hb, hi = i + hb >= n ? n - 1 : i + hb;
        for (int64_t j = lo; j <= hi; j++) m->col[p++] = (int32_t)j;
        m->ptr[i + 1] = p;
    }
    m->nnz = p;
    fill_values_pm1(m, &r);
    fill_x_pm1(m, &r, 1);
    return m;
}

/* Block-diagonal with dense bs x bs blocks (last block truncated). */
vg_csr *vg_blockdiag(int64_t n, int bs, uint64_t seed)
{
    vg_rng r;
    rng_seed(&r, seed);
    vg_csr *m = csr_alloc(n, n, n * (int64_t)bs);
    int64_t p = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t b0 = (i / bs) * bs, b1 = b0 + bs > n ? n : b0 + bs;
        for (int64_t j = b0; j < b1; j++) m->col[p++] = (int32_t)j;
        m->ptr[i + 1] = p;
    }
    m->nnz = p;
    fill_values_pm1(m, &r);
    fill_x_pm1(m, &r, 1);
    return m;
}

/* Uniform random columns, row lengths uniform in [k/2, 3k/2]. */
vg_csr *vg_uniform(int64_t n, int k, uint64_t seed)
{
    vg_rng r;
    rng_seed(&r, seed);
    int64_t *lens = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t lo = k / 2, span = k + 1;
    for (int64_t i = 0; i < n; i++) lens[i] = lo + (int64_t)rng_below(&r, (uint64_t)span);
    vg_csr *m = csr_random_rows(&r, n, n, lens);
    free(lens);
    fill_values_pm1(m, &r);
    fill_x_pm1(m, &r, 1);
    return m;
}

/* Power-law row lengths: len = floor(xmin * U^(-1/(alpha-1))), capped at
 * cap; uniform random columns. */
vg_csr *vg_powerlaw(int64_t n, double alpha, int xmin, int64_t cap, uint64_t seed)
{
    vg_rng r;
    rng_seed(&r, seed);
    int64_t *lens = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; i++) {
        double u = 1.0 - rng_u01(&r); /* (0, 1] */
        double len = floor((double)xmin * pow(u, -1.0 / (alpha - 1.0)));
        if (len > (double)cap) len = (double)cap;
        lens[i] = (int64_t)len;
    }
    vg_csr *m = csr_random_rows(&r, n, n, lens);
    free(lens);
    fill_values_pm1(m, &r);
    fill_x_pm1(m, &r, 1);
    return m;
}

/* Caller-supplied CSR copied into a generator record with seeded values
 * and x (used to wrap the small MatrixMarket fixtures). */
vg_csr *vg_from_pattern(int64_t nrows, int64_t ncols, const int64_t *ptr, const int32_t *col,
                        uint64_t seed)
{
    vg_rng r;
    rng_seed(&r, seed);
    vg_csr *m = csr_alloc(nrows, ncols, ptr[nrows]);
    memcpy(m->ptr, ptr, sizeof(int64_t) * (size_t)(nrows + 1));
    memcpy(m->col, col, sizeof(int32_t) * (size_t)ptr[nrows]);
    fill_values_pm1(m, &r);
    fill_x_pm1(m, &r, 1);
    return m;
}

/* Raw stream access for tests of the generator itself. */
void vg_stream(uint64_t seed, int64_t n, uint64_t *out)
{
    vg_rng r;
    rng_seed(&r, seed);
    for (int64_t i = 0; i < n; i++) out[i] = rng_next(&r);
}
