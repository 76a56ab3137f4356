/*
 * spb_pattern.c — variant G (pattern slices): choosing the dominant row
 * pattern, splitting rows into regular and irregular ones on the GPU,
 * laying out the regular rows' values and building the sub-plan that runs
 * the irregular rows.
 *
 * A row pattern is a list of column offsets P[0 .. L) (ascending, as the
 * CSR keeps columns): row i fits it when each of its columns is i + P[l]
 * for a distinct l.  Stencils and banded matrices have one pattern that
 * every row fits — interior rows hold all of it, boundary rows the part
 * inside the domain — and the plan stores, per row, the mask of positions
 * held plus the values.
 */
#include <stdlib.h>
#include <string.h>

#include "spb_internal.h"

/* Sample rows at multiplicative-hash positions (s + 1) 2654435761 mod
 * 2^32, scaled to [0, nrows): evenly spaced samples would alias with the
 * layout of grid matrices (every sample of a 64^3 stencil at 32 even
 * steps is a boundary row). */
#define PATTERN_SAMPLES 32

/* Offsets of row i into off[]; returns the row length, or -1 when it is
 * empty or longer than VNEC_PATTERN_MAX. */
static int row_offsets(const spb_csr *a, int64_t i, int32_t *off)
{
    int64_t b = (int64_t)a->ptr[i], e = (int64_t)a->ptr[i + 1];
    if (e - b < 1 || e - b > VNEC_PATTERN_MAX) return -1;
    for (int64_t j = b; j < e; j++) off[j - b] = (int32_t)((int64_t)a->indices[j] - i);
    return (int)(e - b);
}

int vnec_pattern_choose(const spb_csr *a, int32_t pat[VNEC_PATTERN_MAX])
{
    int64_t nr = (int64_t)a->base.row;
    if (nr < 1) return 0;
    int32_t off[PATTERN_SAMPLES][VNEC_PATTERN_MAX];
    int len[PATTERN_SAMPLES];
    for (int s = 0; s < PATTERN_SAMPLES; s++) {
        uint32_t h = (uint32_t)((uint64_t)(s + 1) * 2654435761u);
        int64_t i = (int64_t)(((uint64_t)h * (uint64_t)nr) >> 32);
        len[s] = row_offsets(a, i, off[s]);
    }
    int best = -1, best_n = 0;
    for (int s = 0; s < PATTERN_SAMPLES; s++) {
        if (len[s] < 1) continue;
        int n = 0;
        for (int t = 0; t < PATTERN_SAMPLES; t++)
            n += len[t] == len[s] && !memcmp(off[t], off[s], sizeof(int32_t) * (size_t)len[s]);
        if (n > best_n) {
            best_n = n;
            best = s;
        }
    }
    if (best < 0) return 0;
    memcpy(pat, off[best], sizeof(int32_t) * (size_t)len[best]);
    return len[best];
}

/* masks (nrows) and flags (nrows + 1 device ints, the last one 0: flag 1
 * marks a row that does not fit the pattern) for a pattern. */
static SpB_Info match_device(spb_csr *a, const int32_t *pat, int L, uint32_t *masks, int32_t *flags)
{
    hipStream_t s = vnec_stream();
    int64_t nr = (int64_t)a->base.row;
    int32_t *dpat = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)L, NULL);
    SpB_Info rc = dpat ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(dpat, pat, sizeof(int32_t) * (size_t)L, hipMemcpyHostToDevice, s), "h2d");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemsetAsync(flags + nr, 0, sizeof(int32_t), s), "memset");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_pattern_match(a->d_ptr, a->d_col, nr, dpat, L, masks, flags, s),
                         "pattern match");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    vnec_free_dev(dpat);
    return rc;
}

SpB_Info vnec_pattern_coverage(spb_csr *a, int64_t *nregular)
{
    *nregular = 0;
    int32_t pat[VNEC_PATTERN_MAX];
    int L = vnec_pattern_choose(a, pat);
    if (L < 1) return SpB_SUCCESS;
    if (spb_matrix_upload(a) != SpB_SUCCESS) return SpB_FAILURE;
    int64_t nr = (int64_t)a->base.row, nirr = 0;
    int32_t *flags = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nr + 1), NULL);
    uint32_t *masks = (uint32_t *)vnec_malloc_dev(sizeof(uint32_t) * (size_t)(nr ? nr : 1), NULL);
    SpB_Info rc = (flags && masks) ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS) rc = match_device(a, pat, L, masks, flags);
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_scan_excl_i32(flags, flags, nr + 1, &nirr, vnec_stream()), "scan");
    if (rc == SpB_SUCCESS) *nregular = nr - nirr;
    vnec_free_dev(flags);
    vnec_free_dev(masks);
    return rc;
}

SpB_Info vnec_pattern_ext(const SpB_Matrix A, int32_t pattern[VNEC_PATTERN_MAX], int *L,
                          int64_t *nregular)
{
    spb_csr *a = spb_as_csr(A);
    if (!a || !pattern || !L || !nregular) return vnec_fail("vnec_pattern_ext: bad arguments");
    *L = vnec_pattern_choose(a, pattern);
    *nregular = 0;
    if (*L < 1) return SpB_SUCCESS;
    return vnec_pattern_coverage(a, nregular);
}

SpB_Info vnec_pattern_match_ext(const SpB_Matrix A, const int32_t *pattern, int L, uint32_t *masks,
                                int32_t *flags)
{
    spb_csr *a = spb_as_csr(A);
    if (!a || !pattern || !masks || !flags || L < 1 || L > VNEC_PATTERN_MAX)
        return vnec_fail("vnec_pattern_match_ext: bad arguments (1 <= L <= %d)", VNEC_PATTERN_MAX);
    if (spb_matrix_upload(a) != SpB_SUCCESS) return SpB_FAILURE;
    int64_t nr = (int64_t)a->base.row;
    int32_t *d = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nr + 1), NULL);
    uint32_t *dm = (uint32_t *)vnec_malloc_dev(sizeof(uint32_t) * (size_t)(nr ? nr : 1), NULL);
    SpB_Info rc = (d && dm) ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS) rc = match_device(a, pattern, L, dm, d);
    if (rc == SpB_SUCCESS && nr)
        rc = vnec_hip_ok(hipMemcpy(flags, d, sizeof(int32_t) * (size_t)nr, hipMemcpyDeviceToHost), "d2h");
    if (rc == SpB_SUCCESS && nr)
        rc = vnec_hip_ok(hipMemcpy(masks, dm, sizeof(uint32_t) * (size_t)nr, hipMemcpyDeviceToHost), "d2h");
    vnec_free_dev(d);
    vnec_free_dev(dm);
    return rc;
}

/* CSR of the irregular rows (host arrays of a), as a matrix of this
 * library; the sub-plan runs on it. */
static SpB_Info irregular_submatrix(const spb_csr *a, const int32_t *irr, int64_t nirr, SpB_Matrix *out)
{
    const int fp32 = a->base.element_type == SpB_FP32; /* the sub-matrix keeps the value type */
    const size_t es = fp32 ? sizeof(float) : sizeof(double);
    int64_t nnz = 0;
    for (int64_t k = 0; k < nirr; k++) nnz += (int64_t)(a->ptr[irr[k] + 1] - a->ptr[irr[k]]);
    int64_t *ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nirr + 1));
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    char *val = (char *)malloc(es * (size_t)(nnz ? nnz : 1));
    SpB_Info rc = (ptr && col && val) ? SpB_SUCCESS : vnec_fail("pattern plan: out of memory");
    if (rc == SpB_SUCCESS) {
        int64_t z = 0;
        ptr[0] = 0;
        for (int64_t k = 0; k < nirr; k++) {
            for (SpB_Index j = a->ptr[irr[k]]; j < a->ptr[irr[k] + 1]; j++, z++) {
                col[z] = (int32_t)a->indices[j];
                memcpy(val + es * (size_t)z, (const char *)a->val + es * (size_t)j, es);
            }
            ptr[k + 1] = z;
        }
    }
    *out = NULL;
    if (rc == SpB_SUCCESS) rc = SpB_Matrix_new(out, fp32 ? SpB_FP32 : SpB_FP64, (SpB_Index)nirr, a->base.col);
    if (rc == SpB_SUCCESS) {
        ((spb_csr *)*out)->base.format_type = SpB_CSR; /* whatever the global default */
        rc = fp32 ? SpB_Matrix_import_csr_fp32_ext(*out, ptr, col, (const float *)val)
                  : SpB_Matrix_import_csr_ext(*out, ptr, col, (const double *)val);
    }
    if (rc != SpB_SUCCESS && *out) SpB_Matrix_free(out);
    free(ptr);
    free(col);
    free(val);
    return rc;
}

void vnec_pattern_window(const int32_t *pat, int L, vnec_pattern_args *g)
{
    const int R = VNEC_PATTERN_ROWS; /* rows per workgroup of the pattern kernel */
    int nwin = 0, wsize = 0, win_of[VNEC_PATTERN_MAX];
    for (int l = 0; l < L; l++) { /* offsets ascend: merge overlapping or adjacent runs */
        const int64_t lo = pat[l], hi = (int64_t)pat[l] + R; /* [lo, hi) */
        if (nwin > 0 && lo <= (int64_t)g->wst[nwin - 1] + g->wlen[nwin - 1]) {
            const int64_t end = (int64_t)g->wst[nwin - 1] + g->wlen[nwin - 1];
            if (hi > end) g->wlen[nwin - 1] = (int32_t)(hi - g->wst[nwin - 1]);
        } else {
            g->wst[nwin] = pat[l];
            g->wlen[nwin] = R;
            nwin++;
        }
        win_of[l] = nwin - 1;
    }
    for (int k = 0; k < nwin; k++) {
        g->wbase[k] = wsize;
        wsize += g->wlen[k];
    }
    for (int l = 0; l < L; l++) g->wpos[l] = g->wbase[win_of[l]] + (pat[l] - g->wst[win_of[l]]);
    g->nwin = nwin;
    g->wsize = wsize <= VNEC_PATTERN_WIN_MAX ? wsize : 0;
}

SpB_Info vnec_pattern_build(struct vnec_plan_s *p, spb_csr *a)
{
    hipStream_t s = vnec_stream();
    int64_t nr = p->nrows;
    int32_t pat[VNEC_PATTERN_MAX];
    int L = vnec_pattern_choose(a, pat);
    if (L < 1) return vnec_fail("variant G: no row pattern (sample rows are empty or longer than %d)",
                                VNEC_PATTERN_MAX);
    p->g_L = L;
    memcpy(p->g_pat_h, pat, sizeof(int32_t) * (size_t)L);
    p->g_nslices = (nr + 63) / 64;
    p->g_pat = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)L, &p->dev_bytes);
    p->g_masks = (uint32_t *)vnec_malloc_dev(sizeof(uint32_t) * (size_t)(nr ? nr : 1), &p->dev_bytes);
    p->g_vals = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(p->g_nslices ? p->g_nslices * L * 64 : 1),
                                          &p->dev_bytes);
    int32_t *flags = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nr + 1), NULL);
    int32_t *pos = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nr + 1), NULL);
    SpB_Info rc = (p->g_pat && p->g_masks && p->g_vals && flags && pos) ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(p->g_pat, pat, sizeof(int32_t) * (size_t)L, hipMemcpyHostToDevice, s), "h2d");
    if (rc == SpB_SUCCESS) rc = match_device(a, pat, L, p->g_masks, flags);
    int64_t nirr = 0;
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_scan_excl_i32(flags, pos, nr + 1, &nirr, s), "scan");
    p->g_nirr = nirr;
    if (rc == SpB_SUCCESS) {
        p->g_irr = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nirr ? nirr : 1), &p->dev_bytes);
        if (!p->g_irr) rc = SpB_FAILURE;
    }
    if (rc == SpB_SUCCESS && nirr)
        rc = vnec_hip_ok(vnec_hip_pattern_split(flags, pos, nr, p->g_irr, s), "pattern split");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_pattern_fill(a->d_ptr, a->d_val, p->g_masks, nr, L, p->g_vals, s),
                         "pattern fill");
    const int enc = p->tune.encoding;
    if (rc == SpB_SUCCESS && enc != VNEC_ENC_FP64) { /* narrow the slots: hi (+ lo) */
        const int64_t n = p->g_nslices * L * 64;
        p->g_hi = (float *)vnec_malloc_dev(sizeof(float) * (size_t)(n ? n : 1), &p->dev_bytes);
        if (enc == VNEC_ENC_HILO)
            p->g_lo = (float *)vnec_malloc_dev(sizeof(float) * (size_t)(n ? n : 1), &p->dev_bytes);
        rc = (p->g_hi && (enc != VNEC_ENC_HILO || p->g_lo)) ? SpB_SUCCESS : SpB_FAILURE;
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(vnec_hip_split_hilo(p->g_vals, n, p->g_hi, p->g_lo, s), "split slots");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
        if (rc == SpB_SUCCESS) {
            vnec_free_dev(p->g_vals);
            p->g_vals = NULL;
            p->dev_bytes -= (int64_t)(sizeof(double) * (size_t)(n ? n : 1));
        }
    }
    int32_t *hirr = NULL;
    if (rc == SpB_SUCCESS && nirr) {
        hirr = (int32_t *)malloc(sizeof(int32_t) * (size_t)nirr);
        rc = hirr ? SpB_SUCCESS : vnec_fail("pattern plan: out of memory");
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(hipMemcpyAsync(hirr, p->g_irr, sizeof(int32_t) * (size_t)nirr, hipMemcpyDeviceToHost, s), "d2h");
    }
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    vnec_free_dev(flags);
    vnec_free_dev(pos);
    /* rows that do not fit: their own CSR and a plan the learned selector
     * picks for it (explicit tuning: never variant G again) */
    if (rc == SpB_SUCCESS && nirr) rc = irregular_submatrix(a, hirr, nirr, &p->g_sub);
    free(hirr);
    if (rc == SpB_SUCCESS && nirr && enc == VNEC_ENC_FP32) {
        /* FP32 rows: the FP32 matrices' default plan (slices or merge-path) */
        rc = vnec_plan_build_ext(&p->g_subplan, p->g_sub, NULL);
        if (rc == SpB_SUCCESS) {
            p->g_ysub = (double *)vnec_malloc_dev(sizeof(float) * (size_t)nirr, &p->dev_bytes);
            if (!p->g_ysub) rc = SpB_FAILURE;
        }
    } else if (rc == SpB_SUCCESS && nirr) {
        vnec_tuning t;
        rc = vnec_select_tuning_no_pattern(spb_as_csr(p->g_sub), &t);
        if (rc == SpB_SUCCESS && enc != VNEC_ENC_FP64) { /* the same value encoding */
            if (t.variant == VNEC_VAR_D || t.variant == VNEC_VAR_H || t.variant == VNEC_VAR_C) { /* FP64 only */
                vnec_tuning l = {VNEC_VAR_L, 7, 64, enc, 1, 0};
                t = l;
            }
            t.encoding = enc;
        }
        if (rc == SpB_SUCCESS) rc = vnec_plan_build_ext(&p->g_subplan, p->g_sub, &t);
        if (rc == SpB_SUCCESS) {
            p->g_ysub = (double *)vnec_malloc_dev(sizeof(double) * (size_t)nirr, &p->dev_bytes);
            if (!p->g_ysub) rc = SpB_FAILURE;
        }
    }
    return rc;
}
