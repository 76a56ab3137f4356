/*
 * vnec_oracle.c — CPU restatement of the reference VNEC FP64 SpMV path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the shipped library.  Every function names the reference
 * file:line whose behaviour it restates (paths relative to the VNEC repo).
 *
 * Parameters that the reference fixes at compile time (NUM_THREADS = P,
 * nLanes_f64 = L, IRD_thr_fp64, the 32-entry cross-row length cut) are
 * run-time arguments here.  At P = 36, L = 8 every integer structure must be
 * bit-exact with the reference build (pinned against the golden fixtures
 * written from oracle/_ref).
 *
 * Deliberate differences from the reference (its defects, not its design):
 *   - the D carry starts at max(ptr[end.x], start.y), so a partition lying
 *     inside one row is not counted twice (spmv_vnec_D_for_mxv.c:471);
 *   - carries into one row are added in partition order, never raced
 *     (spmv_vnec_D_for_mxv.c:480-484, spmv_vnec_L_for_mxv.c:160-164);
 *   - compact-x reads past NEC use zero padding instead of heap garbage
 *     (spmv_vnec_D_for_mxv.c:449);
 *   - the S row snap clamps x + 1 to nrows (spmv_vnec_S_for_mxv.c:363-370).
 *
 * Parity status of each function (what pins it) is listed in the header of
 * tests/test_oracle.py; nothing here is "parity unpinned".
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { int32_t x, y; } orc_coord;

/* ------------------------------------------------------------------ a1
 * MergePathDivide — src/operation/src/spmv_vnec_common.c:3-33.
 * Merge list A is the row-end offsets ptr[1..nrows], list B the natural
 * numbers 0..nnz-1.  Binary search along the diagonal for the first x with
 * A[x] > B[d - x - 1]; the split is (x, d - x).  Closed form:
 *   x = #{ i : ptr[i+1] + i + 1 <= d }  clipped to [max(d - nnz, 0), min(d, nrows)].
 */
void orc_merge_path(int64_t d, const int64_t *ptr, int64_t nrows, int64_t nnz, orc_coord *out)
{
    int64_t lo = d - nnz > 0 ? d - nnz : 0;
    int64_t hi = d < nrows ? d : nrows;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (ptr[mid + 1] <= d - mid - 1)
            lo = mid + 1;
        else
            hi = mid;
    }
    out->x = (int32_t)lo;
    out->y = (int32_t)(d - lo);
}

/* Diagonals d_t = min(t * ceil(M / P), M), M = nnz + nrows —
 * spmv_vnec_D_for_mxv.c:250-268 (same in S :344-362, L :10-33, IRD :172-190). */
int64_t orc_items_per_part(int64_t nrows, int64_t nnz, int64_t P)
{
    int64_t M = nnz + nrows;
    return (M + P - 1) / P;
}

void orc_partition(const int64_t *ptr, int64_t nrows, int64_t nnz, int64_t P, orc_coord *start,
                   orc_coord *end)
{
    int64_t M = nnz + nrows;
    int64_t ipt = orc_items_per_part(nrows, nnz, P);
    for (int64_t t = 0; t < P; t++) {
        int64_t ds = ipt * t < M ? ipt * t : M;
        int64_t de = ds + ipt < M ? ds + ipt : M;
        orc_merge_path(ds, ptr, nrows, nnz, &start[t]);
        orc_merge_path(de, ptr, nrows, nnz, &end[t]);
    }
}

/* ------------------------------------------------------------------ a2
 * S-variant row snap — spmv_vnec_S_for_mxv.c:356-371.  Partition t owns
 * the whole rows [mp(d_t).x + 1, mp(d_t+1).x + 1); the first starts at row
 * 0, the last ends at nrows; .y = ptr[.x].  x + 1 is clamped to nrows. */
void orc_partition_snap(const int64_t *ptr, int64_t nrows, int64_t nnz, int64_t P,
                        orc_coord *start, orc_coord *end)
{
    orc_partition(ptr, nrows, nnz, P, start, end);
    for (int64_t t = 0; t < P; t++) {
        int64_t sx = start[t].x + 1, ex = end[t].x + 1;
        if (t == 0) sx = 0;
        if (t == P - 1) ex = nrows;
        if (sx > nrows) sx = nrows;
        if (ex > nrows) ex = nrows;
        start[t].x = (int32_t)sx;
        end[t].x = (int32_t)ex;
        start[t].y = (int32_t)ptr[sx];
        end[t].y = (int32_t)ptr[ex];
    }
}

/* ------------------------------------------------------------------ a3
 * ECR column compression — spmv_vnec_D_for_mxv.c:280-338 (S :388-441,
 * IRD ird_calc.c:29-65).  For the nnz range [start.y, end.y) of each
 * partition: win = sorted distinct columns (use_x_indices), nec = |win|
 * (NEC_NUM), ecr[j] = rank of col[j] in win (ecr_indices).
 * Restated as sort + unique + lower_bound.  win is written partition after
 * partition from win[0]; win_off (P + 1 entries) gives each partition's
 * first entry.  win must hold nnz entries (the total is <= nnz). */
static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

void orc_ecr(const int32_t *col, int64_t P, const orc_coord *start, const orc_coord *end,
             int32_t *ecr, int32_t *nec, int64_t *win_off, int32_t *win)
{
    int64_t off = 0;
    for (int64_t t = 0; t < P; t++) {
        int64_t ys = start[t].y, ye = end[t].y;
        int64_t n = ye > ys ? ye - ys : 0;
        int32_t *w = win + off;
        memcpy(w, col + ys, sizeof(int32_t) * (size_t)n);
        qsort(w, (size_t)n, sizeof(int32_t), cmp_i32);
        int64_t u = 0;
        for (int64_t k = 0; k < n; k++)
            if (k == 0 || w[k] != w[u - 1]) w[u++] = w[k];
        for (int64_t j = ys; j < ye; j++) {
            int64_t lo = 0, hi = u;
            while (lo < hi) {
                int64_t mid = (lo + hi) / 2;
                if (w[mid] < col[j]) lo = mid + 1; else hi = mid;
            }
            ecr[j] = (int32_t)lo;
        }
        nec[t] = (int32_t)u;
        win_off[t] = off;
        off += u;
    }
    win_off[P] = off;
}

/* ------------------------------------------------------------------ a4
 * VNEC-D grouping — spmv_vnec_D_for_mxv.c:356-392.  Rows [start.x, end.x)
 * of partition t, each from max(ptr[i], start.y) to ptr[i+1]: a greedy
 * group opens at the first compressed column c0 and takes every following
 * entry with ecr - c0 < L (val_align[g*L + ecr - c0] = a, the last
 * duplicate wins); the next entry opens a new group.  v_row_ptr[i+1] is the
 * running group count.  col_start / val_align may be NULL (count only, as
 * IRD does at ird_calc.c:237-267); val_align must be zeroed by the caller.
 * Returns the total number of groups G. */
int64_t orc_d_build(const int64_t *ptr, const double *val, const int32_t *ecr, int64_t nrows,
                    int64_t P, int64_t L, const orc_coord *start, const orc_coord *end,
                    int32_t *v_row_ptr, int32_t *col_start, double *val_align)
{
    int64_t g = 0;
    if (v_row_ptr) v_row_ptr[0] = 0;
    (void)nrows;
    for (int64_t t = 0; t < P; t++) {
        for (int64_t i = start[t].x; i < end[t].x; i++) {
            int64_t b = ptr[i] > start[t].y ? ptr[i] : start[t].y;
            int64_t e = ptr[i + 1];
            if (e > b) {
                int64_t c0 = ecr[b];
                if (col_start) col_start[g] = (int32_t)c0;
                if (val_align) val_align[g * L] = val[b];
                for (int64_t j = b + 1; j < e; j++) {
                    int64_t dist = ecr[j] - c0;
                    if (dist < L) {
                        if (val_align) val_align[g * L + dist] = val[j];
                    } else {
                        g++;
                        c0 = ecr[j];
                        if (col_start) col_start[g] = (int32_t)c0;
                        if (val_align) val_align[g * L] = val[j];
                    }
                }
                g++;
            }
            if (v_row_ptr) v_row_ptr[i + 1] = (int32_t)g;
        }
    }
    return g;
}

/* ------------------------------------------------------------------ a10
 * IRD_VNEC — src/operation/src/ird_calc.c:2-123.  IRD = (float)nnz /
 * (float)(G * L) in single precision; D when IRD >= (float)thr, else S.
 * Returns 0 (SpB_VNEC_D) or 1 (SpB_VNEC_S); *ird receives the ratio. */
int orc_ird(const int64_t *ptr, const int32_t *col, int64_t nrows, int64_t nnz, int64_t P,
            int64_t L, double thr, float *ird)
{
    orc_coord *s = (orc_coord *)malloc(sizeof(orc_coord) * (size_t)P);
    orc_coord *e = (orc_coord *)malloc(sizeof(orc_coord) * (size_t)P);
    orc_partition(ptr, nrows, nnz, P, s, e);
    int32_t *ecr = (int32_t *)calloc((size_t)(nnz > 0 ? nnz : 1), sizeof(int32_t));
    int32_t *nec = (int32_t *)malloc(sizeof(int32_t) * (size_t)P);
    int64_t *woff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(P + 1));
    int32_t *win = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    orc_ecr(col, P, s, e, ecr, nec, woff, win);
    int64_t G = orc_d_build(ptr, NULL, ecr, nrows, P, L, s, e, NULL, NULL, NULL);
    int32_t slots = (int32_t)(G * L);
    float r = (float)nnz / (float)slots;
    float th = (float)thr;
    if (ird) *ird = r;
    free(s); free(e); free(ecr); free(nec); free(woff); free(win);
    return r >= th ? 0 : 1;
}

/* ---------------------------------------------------------------- a7/a8
 * VNEC-S task order, reorder and transpose — spmv_vnec_S_for_mxv.c:442-585.
 * Per snapped partition: rows bucketed by length ascending, ascending row
 * id within a length.  For a length <= cross_max the first cnt - cnt % L
 * rows go to the cross-row part ("order"), the rest to "remain"; longer
 * rows all go to "remain".  tasks (concatenated over partitions, nrows
 * entries in all) = order || remain per partition; spvv_len[t] = |order|.
 * Rows are then copied in task order (row_begin/row_end, s_val, s_col with
 * the ECR rank); each cross-row group of L equal-length rows is rewritten
 * from [row][l] to [l][row]. */
void orc_s_build(const int64_t *ptr, const double *val, const int32_t *ecr, int64_t nrows,
                 int64_t P, int64_t L, int64_t cross_max, const orc_coord *start,
                 const orc_coord *end, int32_t *tasks, int32_t *spvv_len, int32_t *row_begin,
                 int32_t *row_end, double *s_val, int32_t *s_col, int32_t *task_start,
                 int32_t *task_end)
{
    (void)nrows;
    int64_t tpos = 0;
    for (int64_t t = 0; t < P; t++) {
        int64_t rs = start[t].x, re = end[t].x;
        int64_t nr = re > rs ? re - rs : 0;
        int64_t maxlen = 0;
        for (int64_t r = rs; r < re; r++)
            if (ptr[r + 1] - ptr[r] > maxlen) maxlen = ptr[r + 1] - ptr[r];
        /* counting sort by length, stable in row id */
        int64_t *cnt = (int64_t *)calloc((size_t)maxlen + 2, sizeof(int64_t));
        for (int64_t r = rs; r < re; r++) cnt[ptr[r + 1] - ptr[r] + 1]++;
        int64_t *first = (int64_t *)malloc(sizeof(int64_t) * ((size_t)maxlen + 2));
        memcpy(first, cnt, sizeof(int64_t) * ((size_t)maxlen + 2));
        for (int64_t k = 1; k <= maxlen + 1; k++) first[k] += first[k - 1];
        int32_t *byLen = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nr > 0 ? nr : 1));
        int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * ((size_t)maxlen + 1));
        for (int64_t k = 0; k <= maxlen; k++) fill[k] = first[k];
        for (int64_t r = rs; r < re; r++) byLen[fill[ptr[r + 1] - ptr[r]]++] = (int32_t)r;
        int32_t *out = tasks + tpos;
        int64_t no = 0, nrem = 0;
        /* cross-row part first */
        for (int64_t len = 0; len <= maxlen; len++) {
            int64_t c = cnt[len + 1];
            if (!c) continue;
            int64_t bulk = len > cross_max ? 0 : c - c % L;
            for (int64_t k = 0; k < bulk; k++) out[no++] = byLen[first[len] + k];
        }
        nrem = no;
        for (int64_t len = 0; len <= maxlen; len++) {
            int64_t c = cnt[len + 1];
            if (!c) continue;
            int64_t bulk = len > cross_max ? 0 : c - c % L;
            for (int64_t k = bulk; k < c; k++) out[nrem++] = byLen[first[len] + k];
        }
        spvv_len[t] = (int32_t)no;
        free(cnt); free(first); free(byLen); free(fill);
        tpos += nr;
    }
    /* reorder in task order */
    int64_t pos = 0, npos = 0, tp = 0;
    for (int64_t t = 0; t < P; t++) {
        int64_t nr = end[t].x > start[t].x ? end[t].x - start[t].x : 0;
        task_start[t] = (int32_t)tp;
        task_end[t] = (int32_t)(tp + nr);
        for (int64_t k = 0; k < nr; k++) {
            int64_t row = tasks[tp + k];
            row_begin[pos] = (int32_t)npos;
            for (int64_t j = ptr[row]; j < ptr[row + 1]; j++) {
                if (s_val) s_val[npos] = val[j];
                s_col[npos] = ecr[j];
                npos++;
            }
            row_end[pos] = (int32_t)npos;
            pos++;
        }
        tp += nr;
    }
    /* transpose each cross-row group to [l][lane] */
    for (int64_t t = 0; t < P; t++) {
        int64_t p0 = task_start[t];
        for (int64_t c = 0; c < spvv_len[t]; c += L) {
            int64_t p = p0 + c;
            int64_t len = row_end[p] - row_begin[p];
            int64_t base = row_begin[p];
            if (len == 0) continue;
            int32_t *tc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(len * L));
            double *tv = s_val ? (double *)malloc(sizeof(double) * (size_t)(len * L)) : NULL;
            memcpy(tc, s_col + base, sizeof(int32_t) * (size_t)(len * L));
            if (tv) memcpy(tv, s_val + base, sizeof(double) * (size_t)(len * L));
            for (int64_t l = 0; l < len; l++)
                for (int64_t r = 0; r < L; r++) {
                    s_col[base + l * L + r] = tc[r * len + l];
                    if (tv) s_val[base + l * L + r] = tv[r * len + l];
                }
            free(tc);
            free(tv);
        }
    }
}

/* ------------------------------------------------------------------ a13
 * Serial CSR oracle — src/operation/test/test_spmv_vnec.hpp:57-79:
 * y_i = sum_j a_ij x_j, left to right, FP64. */
void orc_spmv_serial(const int64_t *ptr, const int32_t *col, const double *val, const double *x,
                     int64_t nrows, double *y)
{
    for (int64_t i = 0; i < nrows; i++) {
        double s = 0.0;
        for (int64_t j = ptr[i]; j < ptr[i + 1]; j++) s += val[j] * x[col[j]];
        y[i] = s;
    }
}

/* The reference harness's FP32 serial SpMV (test_spmv_vnec.hpp:34-54):
 * per row, t1 += t2 * t3 in float, the product rounded to float before the
 * add (no contraction into a fused multiply-add). */
__attribute__((optimize("fp-contract=off")))
void orc_spmv_serial_f32(const int64_t *ptr, const int32_t *col, const float *val, const float *x,
                         int64_t nrows, float *y)
{
    for (int64_t i = 0; i < nrows; i++) {
        float t1 = 0.0f;
        for (int64_t j = ptr[i]; j < ptr[i + 1]; j++) {
            float t2 = val[j], t3 = x[col[j]];
            t1 += t2 * t3;
        }
        y[i] = t1;
    }
}

/* Tolerance scale (|A||x|)_i and a long-double row sum for the per-row
 * gate |y - y_ref| <= tol * (|A||x|)_i. */
void orc_spmv_abs(const int64_t *ptr, const int32_t *col, const double *val, const double *x,
                  int64_t nrows, double *ax, double *y_ld)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nrows; i++) {
        long double s = 0.0L, a = 0.0L;
        for (int64_t j = ptr[i]; j < ptr[i + 1]; j++) {
            long double p = (long double)val[j] * (long double)x[col[j]];
            s += p;
            a += fabsl(p);
        }
        if (ax) ax[i] = (double)a;
        if (y_ld) y_ld[i] = (double)s;
    }
}

/* Multi-RHS serial oracle: X is [ncols][k] row-major, Y is [nrows][k]. */
void orc_spmm_serial(const int64_t *ptr, const int32_t *col, const double *val, const double *X,
                     int64_t nrows, int64_t k, double *Y)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nrows; i++) {
        for (int64_t c = 0; c < k; c++) {
            double s = 0.0;
            for (int64_t j = ptr[i]; j < ptr[i + 1]; j++) s += val[j] * X[(int64_t)col[j] * k + c];
            Y[i * k + c] = s;
        }
    }
}

/* ------------------------------------------------------------------ a11
 * VNEC-L merge-based CSR — src/operation/src/spmv_vnec_L_for_mxv.c:124-165.
 * Each partition walks its rows [start.x, end.x) from start.y, writes the
 * rows it completes, and sums its partial last row into a carry; the
 * carries are then added to y[end.x] for t < P - 1, in partition order. */
void orc_spmv_L(const int64_t *ptr, const int32_t *col, const double *val, const double *x,
                int64_t nrows, int64_t P, const orc_coord *start, const orc_coord *end, double *y,
                int nthreads)
{
    (void)nrows;
    double *cv = (double *)malloc(sizeof(double) * (size_t)P);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t t = 0; t < P; t++) {
        int64_t j = start[t].y;
        for (int64_t i = start[t].x; i < end[t].x; i++) {
            double s = 0.0;
            for (; j < ptr[i + 1]; j++) s += val[j] * x[col[j]];
            y[i] = s;
        }
        double c = 0.0;
        for (; j < end[t].y; j++) c += val[j] * x[col[j]];
        cv[t] = c;
    }
    for (int64_t t = 0; t + 1 < P; t++) y[end[t].x] += cv[t];
    free(cv);
}

/* ------------------------------------------------------------------ a5
 * VNEC-D SpMV — spmv_vnec_D_for_mxv.c:400-485.  Per partition: gather the
 * compact x (xc[k] = x[win[k]]), then y_i = sum over the row's groups of
 * sum_l val_align[g*L + l] * xc[col_start[g] + l]; the partial last row
 * [max(ptr[end.x], start.y), end.y) is carried with xc[ecr[j]] and added
 * in partition order.  xc is zero-padded by L entries. */
void orc_spmv_D(const int64_t *ptr, const double *val, const double *x, int64_t nrows, int64_t P,
                int64_t L, const orc_coord *start, const orc_coord *end, const int32_t *ecr,
                const int32_t *nec, const int64_t *win_off, const int32_t *win,
                const int32_t *v_row_ptr, const int32_t *col_start, const double *val_align,
                double *y, int nthreads)
{
    (void)nrows;
    double *cv = (double *)malloc(sizeof(double) * (size_t)P);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t t = 0; t < P; t++) {
        int64_t n = nec[t];
        double *xc = (double *)calloc((size_t)(n + L), sizeof(double));
        const int32_t *w = win + win_off[t];
        for (int64_t k = 0; k < n; k++) xc[k] = x[w[k]];
        for (int64_t i = start[t].x; i < end[t].x; i++) {
            double s = 0.0;
            for (int64_t g = v_row_ptr[i]; g < v_row_ptr[i + 1]; g++) {
                const double *va = val_align + g * L;
                const double *xv = xc + col_start[g];
                for (int64_t l = 0; l < L; l++) s += va[l] * xv[l];
            }
            y[i] = s;
        }
        double c = 0.0;
        int64_t b = ptr[end[t].x < nrows ? end[t].x : nrows];
        if (b < start[t].y) b = start[t].y;
        for (int64_t j = b; j < end[t].y; j++) c += val[j] * xc[ecr[j]];
        cv[t] = c;
        free(xc);
    }
    for (int64_t t = 0; t + 1 < P; t++) y[end[t].x] += cv[t];
    free(cv);
}

/* ------------------------------------------------------------------ a9
 * VNEC-S SpMV — spmv_vnec_S_for_mxv.c:589-660 with the cross-row and
 * in-row micro-kernels of SpB_mxv.h:128-191: cross-row groups of L
 * equal-length rows accumulate lane by lane over the [l][lane] layout and
 * store y[tasks[c + r]]; the remaining rows are plain dot products. */
void orc_spmv_S(const double *x, int64_t P, int64_t L, const int32_t *nec, const int64_t *win_off,
                const int32_t *win, const int32_t *tasks, const int32_t *spvv_len,
                const int32_t *row_begin, const int32_t *row_end, const double *s_val,
                const int32_t *s_col, const int32_t *task_start, const int32_t *task_end,
                double *y, int nthreads)
{
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t t = 0; t < P; t++) {
        int64_t n = nec[t];
        double *xc = (double *)calloc((size_t)(n + 1), sizeof(double));
        const int32_t *w = win + win_off[t];
        for (int64_t k = 0; k < n; k++) xc[k] = x[w[k]];
        int64_t p0 = task_start[t], p1 = task_end[t];
        const int32_t *rows = tasks + p0;
        int64_t p = p0, c = 0;
        double acc[256];
        for (; c < spvv_len[t]; c += L, p += L) {
            int64_t len = row_end[p] - row_begin[p], base = row_begin[p];
            for (int64_t r = 0; r < L; r++) acc[r] = 0.0;
            for (int64_t l = 0; l < len; l++)
                for (int64_t r = 0; r < L; r++)
                    acc[r] += s_val[base + l * L + r] * xc[s_col[base + l * L + r]];
            for (int64_t r = 0; r < L; r++) y[rows[c + r]] = acc[r];
        }
        for (; p < p1; p++) {
            double s = 0.0;
            for (int64_t j = row_begin[p]; j < row_end[p]; j++) s += s_val[j] * xc[s_col[j]];
            y[rows[p - p0]] = s;
        }
        free(xc);
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
