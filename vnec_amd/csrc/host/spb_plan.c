/*
 * spb_plan.c — GPU execution plans: tiles from the GPU merge-path
 * partitioner, the per-tile ECR windows (variant E), value encodings, the
 * launch sequence (SpMV kernel + carry fix-up) and event timing.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spb_internal.h"

#define TILE_THREADS 256

/* L and C plans: relabelled columns, and C's wide (1024-thread) tiles. */
static int plan_relabelled(const vnec_tuning *t)
{
    return (t->variant == VNEC_VAR_L || t->variant == VNEC_VAR_C) && t->lanes != 64 &&
           (t->lanes & VNEC_LANES_RELABEL);
}
static int plan_block(const vnec_tuning *t)
{
    return t->variant == VNEC_VAR_C && t->lanes != 64 && (t->lanes & VNEC_LANES_WIDE) ? VNEC_WIDE_BLOCK : 0;
}
static int plan_threads(const vnec_tuning *t) { return plan_block(t) ? VNEC_WIDE_BLOCK : TILE_THREADS; }
static int plan_gather(const vnec_tuning *t)
{
    return t->variant == VNEC_VAR_C && t->lanes != 64 && (t->lanes & VNEC_LANES_GATHER) ? 1 : 0;
}

static int valid_ipt(int ipt) { return ipt >= 3 && ipt <= 15 && (ipt & 1); }
/* Tiles per resident workgroup from which per-tile-carry plans use the
 * strided persistent launch (the same tiles and carries, so the same bits):
 * from one tile per resident workgroup on it was 1-8 % faster on R-MAT,
 * uniform and power-law matrices; on a fifth of a tile per workgroup the
 * one-workgroup-per-tile launch was up to 4 % faster. */
#define STRIDED_MIN_TILES_PER_WG 1
/* The largest matrices the compiled selector model was trained on. */
#define SELECTOR_TRAINED_MAX_NNZ ((int64_t)1 << 24)
static SpB_Info build_csort(struct vnec_plan_s *p, spb_csr *a);

static void plan_release(struct vnec_plan_s *p)
{
    if (p->gexec) hipGraphExecDestroy(p->gexec);
    for (int i = 0; i < VNEC_PLAN_GRAPHS; i++)
        if (p->sgraph[i].exec) hipGraphExecDestroy(p->sgraph[i].exec);
    if (p->g_subplan) vnec_plan_free_ext(&p->g_subplan);
    if (p->g_sub) SpB_Matrix_free(&p->g_sub);
    vnec_free_dev(p->g_pat);
    vnec_free_dev(p->g_masks);
    vnec_free_dev(p->g_vals);
    vnec_free_dev(p->g_hi);
    vnec_free_dev(p->g_lo);
    vnec_free_dev(p->xt);
    vnec_free_dev(p->yt);
    vnec_free_dev(p->cperm);
    vnec_free_dev(p->rcol);
    vnec_free_dev(p->xperm);
    vnec_free_dev(p->cs_val);
    vnec_free_dev(p->cs_col);
    vnec_free_dev(p->cs_pos);
    vnec_free_dev(p->cs_base);
    vnec_free_dev(p->g_irr);
    vnec_free_dev(p->g_ysub);
    vnec_free_dev(p->coords);
    vnec_free_dev(p->carry);
    vnec_free_dev(p->wg_row);
    vnec_free_dev(p->wg_carry);
    vnec_free_dev(p->lcol);
    vnec_free_dev(p->run_off);
    vnec_free_dev(p->run_start);
    vnec_free_dev(p->run_wpos);
    vnec_free_dev(p->nec);
    vnec_free_dev(p->hi);
    vnec_free_dev(p->lo);
    vnec_free_dev(p->slice_len);
    vnec_free_dev(p->slice_ptr);
    vnec_free_dev(p->tile_slices);
    vnec_free_dev(p->sv);
    vnec_free_dev(p->shi);
    vnec_free_dev(p->slo);
    vnec_free_dev(p->scol);
    vnec_free_dev(p->slcol);
    vnec_free_dev(p->vrp);
    vnec_free_dev(p->col_start);
    vnec_free_dev(p->dvals);
    vnec_free_dev(p->dvals32);
    vnec_free_dev(p->hcol);
    vnec_free_dev(p->hot_order);
    vnec_free_dev(p->xhot);
    memset(p, 0, sizeof *p);
}

SpB_Info vnec_plan_free_ext(vnec_plan *plan)
{
    if (!plan || !*plan) return vnec_fail("vnec_plan_free_ext: NULL plan");
    struct vnec_plan_s *p = *plan;
    if (p->magic != SPB_MAGIC_PLAN) return vnec_fail("vnec_plan_free_ext: not a plan");
    hipStreamSynchronize(vnec_stream());
    plan_release(p);
    free(p);
    *plan = NULL;
    return SpB_SUCCESS;
}

/* Per-tile ECR windows for variants E and D; *max_runs receives the
 * largest run count of any tile; ecr32 (nnz, may be NULL) receives the
 * tile-local ranks as int32. */
static SpB_Info build_windows(struct vnec_plan_s *p, const spb_csr *a, int32_t *max_runs,
                              int32_t *ecr32)
{
    hipStream_t s = vnec_stream();
    int64_t P = p->ntiles, nnz = p->nnz;
    int32_t mx = 0;
    if (vnec_hip_ok(vnec_hip_max_tile_nnz(p->coords, P, &mx, s), "max tile nnz")) return SpB_FAILURE;
    if (mx > VNEC_ECR_TILE_MAX) return vnec_fail("tile holds %d nonzeros (> %d)", mx, VNEC_ECR_TILE_MAX);
    int32_t *uniq = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nnz ? nnz : 1), NULL);
    int32_t *nruns = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(P + 1), NULL);
    p->lcol = (uint16_t *)vnec_malloc_dev(sizeof(uint16_t) * (size_t)(nnz ? nnz : 1), &p->dev_bytes);
    p->nec = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)P, &p->dev_bytes);
    p->run_off = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(P + 1), &p->dev_bytes);
    SpB_Info rc = (uniq && nruns && p->lcol && p->nec && p->run_off) ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_ecr_tiles_pass1(a->d_col, p->coords, P, p->lcol, ecr32, uniq,
                                                  p->nec, nruns, s), "ecr pass 1");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemsetAsync(nruns + P, 0, sizeof(int32_t), s), "memset");
    *max_runs = 0;
    if (rc == SpB_SUCCESS) {
        int32_t *h = (int32_t *)malloc(sizeof(int32_t) * (size_t)P);
        if (!h) rc = vnec_fail("out of memory");
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(hipMemcpyAsync(h, nruns, sizeof(int32_t) * (size_t)P, hipMemcpyDeviceToHost, s), "d2h");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
        if (rc == SpB_SUCCESS)
            for (int64_t t = 0; t < P; t++)
                if (h[t] > *max_runs) *max_runs = h[t];
        free(h);
    }
    int64_t total = 0;
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_scan_excl_i32(nruns, p->run_off, P + 1, &total, s), "run scan");
    if (rc == SpB_SUCCESS) {
        p->run_start = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(total ? total : 1), &p->dev_bytes);
        p->run_wpos = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(total ? total : 1), &p->dev_bytes);
        if (!p->run_start || !p->run_wpos) rc = SpB_FAILURE;
    }
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_ecr_tiles_pass2(p->coords, P, uniq, p->nec, p->run_off,
                                                  p->run_start, p->run_wpos, s), "ecr pass 2");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "windows sync");
    p->nruns = total;
    vnec_free_dev(uniq);
    vnec_free_dev(nruns);
    return rc;
}

/* Largest x window (entries) a variant-S tile may stage in LDS. */
#define S_WIN_MAX 6144
/* Shortest mean run length (window entries per run) for the S window. */
#define S_RUN_MIN 16

/* Variant S: slices, tiles of whole slices, windows or direct columns. */
static SpB_Info build_slices(struct vnec_plan_s *p, const spb_csr *a)
{
    hipStream_t s = vnec_stream();
    int64_t nr = p->nrows, nnz = p->nnz;
    int64_t ns = (nr + 63) / 64;
    p->nslices = ns;
    SpB_Info rc = SpB_SUCCESS;
    p->slice_len = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(ns + 1), &p->dev_bytes);
    p->slice_ptr = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(ns + 1), &p->dev_bytes);
    if (!p->slice_len || !p->slice_ptr) return SpB_FAILURE;
    rc = vnec_hip_ok(vnec_hip_slice_len(a->d_ptr, nr, ns, p->slice_len, p->slice_ptr, s), "slice len");
    int64_t padded = 0;
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_scan_excl_i32(p->slice_ptr, p->slice_ptr, ns + 1, &padded, s), "slice scan");
    if (rc != SpB_SUCCESS) return rc;
    if (padded + ns >= INT32_MAX) return vnec_fail("padded slice layout exceeds 2^31 entries");
    p->s_padded = padded;
    /* tiles of whole slices, balanced over the padded layout */
    int64_t M = ns + padded;
    int64_t T = 2048 * (int64_t)p->tune.ipt;
    p->ntiles = M > 0 ? (M + T - 1) / T : 0;
    if (p->ntiles == 0) return SpB_SUCCESS;
    p->tile_slices = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * 2 * (size_t)(p->ntiles + 1), &p->dev_bytes);
    p->coords = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * 2 * (size_t)(p->ntiles + 1), &p->dev_bytes);
    if (!p->tile_slices || !p->coords) return SpB_FAILURE;
    rc = vnec_hip_ok(vnec_hip_partition(p->slice_ptr, ns, padded, p->ntiles, 1, p->tile_slices, s), "slice partition");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_slice_tile_rows(p->tile_slices, p->ntiles, a->d_ptr, nr, p->coords, s), "tile rows");
    if (rc != SpB_SUCCESS) return rc;
    /* windows: only when every tile's window is small and made of long runs */
    int32_t mx = 0;
    rc = vnec_hip_ok(vnec_hip_max_tile_nnz(p->coords, p->ntiles, &mx, s), "max tile nnz");
    if (rc != SpB_SUCCESS) return rc;
    uint16_t *lcol = NULL;
    p->win = 0;
    if (mx <= VNEC_ECR_TILE_MAX_BIG && nnz > 0) {
        int64_t P = p->ntiles;
        int32_t *uniq = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)nnz, NULL);
        int32_t *nruns = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(P + 1), NULL);
        int32_t *nec = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(P + 1), NULL);
        lcol = (uint16_t *)vnec_malloc_dev(sizeof(uint16_t) * (size_t)nnz, NULL);
        int32_t *hn = (int32_t *)malloc(sizeof(int32_t) * (size_t)P);
        int32_t *hr = (int32_t *)malloc(sizeof(int32_t) * (size_t)P);
        if (!uniq || !nruns || !nec || !lcol || !hn || !hr) rc = vnec_fail("out of memory");
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(vnec_hip_ecr_tiles_pass1(a->d_col, p->coords, P, lcol, NULL, uniq, nec, nruns, s), "ecr");
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(hipMemcpyAsync(hn, nec, sizeof(int32_t) * (size_t)P, hipMemcpyDeviceToHost, s), "d2h");
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(hipMemcpyAsync(hr, nruns, sizeof(int32_t) * (size_t)P, hipMemcpyDeviceToHost, s), "d2h");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
        int ok = rc == SpB_SUCCESS;
        int32_t mn = 0;
        for (int64_t t = 0; ok && t < P; t++) {
            if (hn[t] > mn) mn = hn[t];
            if (hn[t] > S_WIN_MAX || (int64_t)hn[t] < (int64_t)S_RUN_MIN * hr[t]) ok = 0;
        }
        if (ok) {
            int64_t total = 0;
            p->max_nec = mn;
            p->nec = nec;
            nec = NULL;
            p->run_off = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(P + 1), &p->dev_bytes);
            rc = p->run_off ? SpB_SUCCESS : SpB_FAILURE;
            if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemsetAsync(nruns + P, 0, sizeof(int32_t), s), "memset");
            if (rc == SpB_SUCCESS)
                rc = vnec_hip_ok(vnec_hip_scan_excl_i32(nruns, p->run_off, P + 1, &total, s), "run scan");
            if (rc == SpB_SUCCESS) {
                p->run_start = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(total ? total : 1), &p->dev_bytes);
                p->run_wpos = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(total ? total : 1), &p->dev_bytes);
                if (!p->run_start || !p->run_wpos) rc = SpB_FAILURE;
            }
            if (rc == SpB_SUCCESS)
                rc = vnec_hip_ok(vnec_hip_ecr_tiles_pass2(p->coords, P, uniq, p->nec, p->run_off,
                                                          p->run_start, p->run_wpos, s), "ecr pass 2");
            if (rc == SpB_SUCCESS) {
                p->win = 1;
                p->nruns = total;
            }
        }
        vnec_free_dev(uniq);
        vnec_free_dev(nruns);
        vnec_free_dev(nec);
        free(hn);
        free(hr);
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    }
    /* transposed, padded slice arrays */
    size_t np = (size_t)(padded ? padded : 1);
    int enc = p->tune.encoding;
    if (rc == SpB_SUCCESS) {
        if (enc == VNEC_ENC_FP64) p->sv = (double *)vnec_malloc_dev(sizeof(double) * np, &p->dev_bytes);
        else p->shi = (float *)vnec_malloc_dev(sizeof(float) * np, &p->dev_bytes);
        if (enc == VNEC_ENC_HILO) p->slo = (float *)vnec_malloc_dev(sizeof(float) * np, &p->dev_bytes);
        if (p->win) p->slcol = (uint16_t *)vnec_malloc_dev(sizeof(uint16_t) * np, &p->dev_bytes);
        else p->scol = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * np, &p->dev_bytes);
        if ((enc == VNEC_ENC_FP64 && !p->sv) || (enc != VNEC_ENC_FP64 && !p->shi) ||
            (enc == VNEC_ENC_HILO && !p->slo) || (p->win ? !p->slcol : !p->scol))
            rc = SpB_FAILURE;
    }
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_slice_fill(a->d_ptr, a->d_col, p->win ? lcol : NULL, a->d_val, nr, ns,
                                             p->slice_len, p->slice_ptr, p->sv, p->shi, p->slo,
                                             p->scol, p->slcol, s), "slice fill");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    vnec_free_dev(lcol);
    return rc;
}

/* Variant D: windows, then the D layout of every tile (greedy groups of
 * `lanes` window positions per row segment), re-laid for wave64. */
static SpB_Info build_d(struct vnec_plan_s *p, const spb_csr *a)
{
    hipStream_t s = vnec_stream();
    int64_t nr = p->nrows, nnz = p->nnz, P = p->ntiles;
    int L = p->tune.lanes;
    int32_t max_runs = 0;
    int32_t *ecr = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nnz ? nnz : 1), NULL);
    int32_t *cnt = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nr + 2), NULL);
    int32_t *unsorted = cnt ? cnt + nr + 1 : NULL, h_unsorted = 0;
    p->vrp = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nr + 1), &p->dev_bytes);
    SpB_Info rc = (ecr && cnt && p->vrp) ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS) rc = build_windows(p, a, &max_runs, ecr);
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemsetAsync(unsorted, 0, sizeof(int32_t), s), "memset");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_d_count(a->d_ptr, ecr, nr, p->coords, P, L, cnt, unsorted, s), "d count");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(&h_unsorted, unsorted, sizeof(int32_t), hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    if (rc == SpB_SUCCESS && h_unsorted)
        rc = vnec_fail("variant D needs every row's columns in ascending order (as SpB_Matrix_build_FP64 leaves them)");
    int64_t G = 0;
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_scan_excl_i32(cnt, p->vrp, nr + 1, &G, s), "d scan");
    if (rc == SpB_SUCCESS && G * L >= INT32_MAX) rc = vnec_fail("variant D: %lld groups too many", (long long)G);
    double *va = NULL;
    int64_t nblk = (G + 63) / 64;
    const int fp32 = p->tune.encoding == VNEC_ENC_FP32;
    const size_t esz = fp32 ? sizeof(float) : sizeof(double), nslot = (size_t)(nblk ? nblk * 64 * L : 1);
    if (rc == SpB_SUCCESS) {
        p->groups = G;
        p->col_start = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(G ? G : 1), &p->dev_bytes);
        if (fp32) p->dvals32 = (float *)vnec_malloc_dev(sizeof(float) * nslot, &p->dev_bytes);
        else p->dvals = (double *)vnec_malloc_dev(sizeof(double) * nslot, &p->dev_bytes);
        va = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(G ? G * L : 1), NULL);
        if (!p->col_start || !(fp32 ? (void *)p->dvals32 : (void *)p->dvals) || !va) rc = SpB_FAILURE;
    }
    if (rc == SpB_SUCCESS && G)
        rc = vnec_hip_ok(hipMemsetAsync(va, 0, sizeof(double) * (size_t)(G * L), s), "memset");
    if (rc == SpB_SUCCESS && nblk)
        rc = vnec_hip_ok(hipMemsetAsync(fp32 ? (void *)p->dvals32 : (void *)p->dvals, 0, esz * nslot, s),
                         "memset");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_d_fill(a->d_ptr, ecr, a->d_val, nr, p->coords, P, L, p->vrp,
                                         p->col_start, va, 1, s), "d fill");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(fp32 ? vnec_hip_d_interleave_f32(va, G, L, p->dvals32, s)
                              : vnec_hip_d_interleave(va, G, L, p->dvals, s), "d interleave");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    vnec_free_dev(va);
    vnec_free_dev(ecr);
    vnec_free_dev(cnt);
    return rc;
}

/* Variant H: the hot-column set.  A column earns an LDS slot when it is
 * used more often than the cost of loading the slot into every persistent
 * workgroup (workgroups / 16 requests: 16 entries per 128-byte line). */
static SpB_Info build_hot(struct vnec_plan_s *p, const spb_csr *a)
{
    hipStream_t s = vnec_stream();
    int64_t nc = p->ncols, nnz = p->nnz, H = p->tune.lanes;
    if (H > nc) H = nc;
    int32_t *deg = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nc ? nc : 1), NULL);
    int32_t *slot = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nc ? nc : 1), NULL);
    int32_t *hdeg = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nc ? nc : 1));
    int32_t *hord = (int32_t *)malloc(sizeof(int32_t) * (size_t)(H ? H : 1));
    p->hot_order = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nc ? nc : 1), &p->dev_bytes);
    p->hcol = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nnz ? nnz : 1), &p->dev_bytes);
    SpB_Info rc = (deg && slot && hdeg && hord && p->hot_order && p->hcol) ? SpB_SUCCESS
                                                                           : vnec_fail("out of memory");
    if (rc == SpB_SUCCESS && nc)
        rc = vnec_hip_ok(hipMemsetAsync(deg, 0, sizeof(int32_t) * (size_t)nc, s), "memset");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_col_degree(a->d_col, nnz, deg, s), "column uses");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_hot_order(deg, nc, p->hot_order, s), "hot order");
    if (rc == SpB_SUCCESS && nc)
        rc = vnec_hip_ok(hipMemcpyAsync(hdeg, deg, sizeof(int32_t) * (size_t)nc, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS && H)
        rc = vnec_hip_ok(hipMemcpyAsync(hord, p->hot_order, sizeof(int32_t) * (size_t)H, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    int64_t wgs = p->tune.grid, nhot = 0, thr = wgs / 16 > 1 ? wgs / 16 : 1;
    while (rc == SpB_SUCCESS && nhot < H && hdeg[hord[nhot]] > thr) nhot++;
    p->nhot = (int)nhot;
    if (rc == SpB_SUCCESS) {
        p->xhot = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(nhot ? nhot : 1), &p->dev_bytes);
        if (!p->xhot) rc = SpB_FAILURE;
    }
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_hot_encode(a->d_col, nnz, p->hot_order, nhot, nc, slot, p->hcol, s),
                         "hot encode");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    vnec_free_dev(deg);
    vnec_free_dev(slot);
    free(hdeg);
    free(hord);
    return rc;
}

/* Hot columns a relabelled plan packs first: as many x entries as one
 * XCD's 4 MiB L2 holds. */
#define RELABEL_HOT ((int64_t)1 << 19)

/* Relabelled columns: the RELABEL_HOT most used columns first, by
 * descending use (ties by column), then the others in ascending column
 * order (so the x permutation reads them in address order); the plan's own
 * copy of the columns in the new labels, and the permuted-x scratch.  Rows
 * and the order of each row's entries are unchanged, so the sums are L's,
 * bit for bit. */
static SpB_Info build_relabel(struct vnec_plan_s *p, const spb_csr *a)
{
    hipStream_t s = vnec_stream();
    const int64_t nc = p->ncols, nnz = p->nnz;
    int32_t *deg = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nc ? nc : 1), NULL);
    int32_t *full = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nc ? nc : 1), NULL);
    int32_t *rank = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nc ? nc : 1), NULL);
    p->cperm = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nc ? nc : 1), &p->dev_bytes);
    p->rcol = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nnz ? nnz : 1), &p->dev_bytes);
    p->xperm = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(nc ? nc : 1), &p->dev_bytes);
    SpB_Info rc = (deg && full && rank && p->cperm && p->rcol && p->xperm) ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS && nc) rc = vnec_hip_ok(hipMemsetAsync(deg, 0, sizeof(int32_t) * (size_t)nc, s), "memset");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_col_degree(a->d_col, nnz, deg, s), "column uses");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_hot_order(deg, nc, full, s), "column order");
    if (rc == SpB_SUCCESS) /* deg is free scratch from here */
        rc = vnec_hip_ok(vnec_hip_pack_order(full, nc, nc < RELABEL_HOT ? nc : RELABEL_HOT, rank, deg, p->cperm, s),
                         "hot-first order");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_relabel_cols(a->d_col, nnz, p->cperm, nc, rank, p->rcol, s), "relabel");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    vnec_free_dev(deg);
    vnec_free_dev(full);
    vnec_free_dev(rank);
    if (rc == SpB_SUCCESS) p->col = p->rcol; /* variant L reads them; C sorts them */
    return rc;
}

/* Selection rule for variant G: at least 90 % of the rows fit the
 * dominant pattern, its slots hold at least 80 % nonzeros (little zero
 * padding), and the matrix is large enough (2^20 nonzeros) for the bytes
 * saved to outweigh a possible sub-plan's extra launches. */
#define G_MIN_NNZ (1 << 20)
/* Skewed column use: at most this many distinct columns per nonzero over
 * merge-path windows of 16384 items (R-MAT graphs: 0.45-0.78; uniform and
 * power-law-row matrices: 0.97-1.0), scattered over at least this many runs
 * of consecutive columns per nonzero (R-MAT: 0.29-0.65; banded matrices and
 * stencils, whose tiles already gather contiguous columns: 0.00), on at
 * least 2^20 nonzeros (below, the sorted staging costs more than the
 * gathers save). */
#define CSORT_MAX_DISTINCT 0.8
#define CSORT_MIN_RUNS 0.2
#define CSORT_MIN_NNZ (1 << 20)

/* A merge-path plan on a matrix whose tiles reuse columns stores each tile
 * in column order (variant C): the wavefront's gathers then share cache
 * lines among the most used columns. */
int vnec_csort_rule(spb_csr *a)
{
    uint64_t v[VNEC_NFEAT_INTS];
    if (vnec_feature_ints_ext((SpB_Matrix)a, v) != SpB_SUCCESS) {
        vnec_clear_error(); /* no statistics: keep the selector's plan */
        return 0;
    }
    const double z = (double)v[VF_NNZ];
    return v[VF_NNZ] >= CSORT_MIN_NNZ && (double)v[VF_NEC_LARGE] <= CSORT_MAX_DISTINCT * z &&
           (double)v[VF_RUNS_LARGE] >= CSORT_MIN_RUNS * z;
}

int vnec_pattern_rule(spb_csr *a)
{
    if ((int64_t)a->nnz < G_MIN_NNZ) return 0;
    int64_t nreg = 0;
    if (vnec_pattern_coverage(a, &nreg) != SpB_SUCCESS) {
        vnec_clear_error(); /* no pattern plan: the selector decides */
        return 0;
    }
    int32_t pat[VNEC_PATTERN_MAX];
    int L = vnec_pattern_choose(a, pat);
    double slots = (double)L * (double)a->base.row;
    return (double)nreg >= 0.9 * (double)a->base.row && (double)a->nnz >= 0.8 * slots;
}

SpB_Info vnec_plan_build_ext(vnec_plan *plan, const SpB_Matrix A, const vnec_tuning *t)
{
    if (!plan) return vnec_fail("vnec_plan_build_ext: NULL plan handle");
    *plan = NULL;
    spb_csr *a = spb_as_csr(A);
    if (!a) return SpB_FAILURE;
    if (!vnec_runtime_ready()) return SpB_FAILURE;
    vnec_tuning tune;
    const int fp32 = a->base.element_type == SpB_FP32;
    if (fp32 && t && t->variant == VNEC_SELECT_AUTOTUNE) {
        if (vnec_autotune_fp32(a, &tune) != SpB_SUCCESS) return SpB_FAILURE;
    } else if (fp32 && (!t || t->variant < 0)) {
        /* FP32 matrices (the selector's measurements are of the FP64
         * kernels): pattern slices where the pattern rule holds, else the
         * 64-row slice layout when its padding is small, else merge-path
         * tiles of 7 items per thread with per-tile carries, stored in
         * column order where the column-reuse rule holds */
        vnec_tuning f = {VNEC_VAR_L, 7, 64, VNEC_ENC_FP32, 1, -1};
        if (vnec_pattern_rule(a)) {
            f.variant = VNEC_VAR_G;
            f.ipt = 8;
        } else if (vnec_slice_padding(a) <= 1.15) {
            f.variant = VNEC_VAR_S;
            f.ipt = 3;
        } else if (vnec_csort_rule(a)) {
            f.variant = VNEC_VAR_C;
        }
        tune = f;
    } else if (t && t->variant >= 0) {
        tune = *t;
    } else if (t && t->variant == VNEC_SELECT_RULE && t->encoding == VNEC_ENC_FP64 && t->nrhs <= 1 &&
               vnec_pattern_rule(a)) {
        /* the rule mode: most rows follow one column-offset pattern, so
         * pattern slices move 8 bytes per regular nonzero, fewer than any
         * column-indexed candidate (the learned mode has pattern slices
         * among its candidates instead) */
        vnec_tuning g = {VNEC_VAR_G, 8, 64, VNEC_ENC_FP64, 1, -1};
        tune = g;
    } else {
        int nrhs = t && t->nrhs > 0 ? t->nrhs : 1;
        if (vnec_select_tuning(a, t ? t->variant : VNEC_SELECT_LEARNED, nrhs, &tune) != SpB_SUCCESS)
            return SpB_FAILURE;
        if (t) tune.encoding = t->encoding;
        /* a candidate without a kernel for the requested encoding falls back
         * to the merge-path kernel, which has them all */
        const int ok = tune.variant == VNEC_VAR_S || tune.variant == VNEC_VAR_G ? 1
                       : tune.variant == VNEC_VAR_D ? tune.encoding == VNEC_ENC_FP64
                       : vnec_hip_merge_supported(tune.variant, tune.ipt, tune.encoding, tune.nrhs);
        if (!ok) {
            tune.variant = VNEC_VAR_L;
            tune.ipt = 7;
            tune.lanes = 64;
        }
        const int mode = t ? t->variant : VNEC_SELECT_LEARNED;
        if ((tune.variant == VNEC_VAR_L || tune.variant == VNEC_VAR_C) && tune.nrhs == 1 &&
            tune.encoding == VNEC_ENC_FP64) {
            /* CSR-ordered or column-sorted tiles, persistent or with
             * per-tile carries, at the chosen granularity: the rule mode
             * asks the column-reuse rule; the autotune mode, and the
             * learned mode beyond the matrix sizes its model was trained
             * on, measure the four forms */
            if (mode == VNEC_SELECT_RULE) {
                tune.variant = vnec_csort_rule(a) ? VNEC_VAR_C : VNEC_VAR_L;
                tune.lanes = 64;
            } else if (mode == VNEC_SELECT_AUTOTUNE ||
                       (mode == VNEC_SELECT_LEARNED && (int64_t)a->nnz > SELECTOR_TRAINED_MAX_NNZ)) {
                if (vnec_refine_merge_plan(a, &tune) != SpB_SUCCESS) return SpB_FAILURE;
            }
        }
        if ((tune.variant == VNEC_VAR_S || tune.variant == VNEC_VAR_G) && tune.nrhs == 1 &&
            tune.encoding == VNEC_ENC_FP64 && mode == VNEC_SELECT_LEARNED &&
            (int64_t)a->nnz > SELECTOR_TRAINED_MAX_NNZ) {
            /* beyond the matrix sizes the model was trained on, a pick
             * among the slice layouts is measured: 64-row slices or
             * pattern slices, whichever is faster */
            if (vnec_refine_slice_plan(a, &tune) != SpB_SUCCESS) return SpB_FAILURE;
        }
    }
    if (tune.nrhs <= 0) tune.nrhs = 1;
    if (tune.variant < 0 || tune.variant >= VNEC_VAR_N)
        return vnec_fail("variant %d is not available", tune.variant);
    if (tune.encoding < 0 || tune.encoding >= VNEC_ENC_N) return vnec_fail("bad encoding");
    if (tune.grid < -2) return vnec_fail("grid %d: use -2, -1, 0 or a positive multiple of 8", tune.grid);
    if (tune.grid == -2 && !((tune.variant == VNEC_VAR_L || tune.variant == VNEC_VAR_C) && tune.nrhs <= 1))
        return vnec_fail("grid -2 (strided persistent): variants L and C with one right-hand side");
    if (fp32 != (tune.encoding == VNEC_ENC_FP32))
        return vnec_fail(fp32 ? "an FP32 matrix takes encoding VNEC_ENC_FP32"
                              : "encoding VNEC_ENC_FP32 needs a matrix made with SpB_FP32");
    if (fp32) {
        if (tune.variant != VNEC_VAR_L && tune.variant != VNEC_VAR_C && tune.variant != VNEC_VAR_S &&
            tune.variant != VNEC_VAR_D && tune.variant != VNEC_VAR_G)
            return vnec_fail("encoding VNEC_ENC_FP32: variants L, C, S, D and G");
        if (tune.nrhs != 1) return vnec_fail("encoding VNEC_ENC_FP32 takes one right-hand side");
        if ((tune.variant == VNEC_VAR_L || tune.variant == VNEC_VAR_C) && tune.grid >= 0)
            tune.grid = -1; /* per-tile carries */
    }
    const int merge_form = tune.variant == VNEC_VAR_L || tune.variant == VNEC_VAR_C;
    if (merge_form && tune.lanes != 64 &&
        (tune.lanes & ~(VNEC_LANES_RELABEL | VNEC_LANES_WIDE | VNEC_LANES_GATHER)))
        return vnec_fail("lanes %d: variants L and C take 64 or the flags VNEC_LANES_RELABEL | VNEC_LANES_WIDE | "
                         "VNEC_LANES_GATHER", tune.lanes);
    if ((tune.lanes == VNEC_LANES_RELABEL || (merge_form && tune.lanes != 64 && (tune.lanes & VNEC_LANES_RELABEL))) &&
        !(merge_form && tune.encoding == VNEC_ENC_FP64 && tune.nrhs == 1))
        return vnec_fail("relabelled columns (lanes %d): variants L and C, FP64, one right-hand side",
                         VNEC_LANES_RELABEL);
    /* FP64 wide tiles run persistent or with per-tile carries, FP32 ones
     * (grid -1 above) with per-tile carries */
    if (merge_form && tune.lanes != 64 && (tune.lanes & VNEC_LANES_WIDE) &&
        !(tune.variant == VNEC_VAR_C && tune.nrhs == 1 &&
          (tune.grid < 0 ? vnec_hip_strided_supported(tune.variant, tune.ipt, tune.encoding, VNEC_WIDE_BLOCK, 0)
                         : vnec_hip_stream_supported(tune.variant, tune.ipt, tune.encoding, VNEC_WIDE_BLOCK))))
        return vnec_fail("wide tiles (lanes %d): variant C, FP64 or FP32, one right-hand side, ipt 3, 5, 7, 9 "
                         "or 11", VNEC_LANES_WIDE);
    if (merge_form && tune.lanes != 64 && (tune.lanes & VNEC_LANES_GATHER) &&
        !(tune.variant == VNEC_VAR_C && tune.nrhs == 1 && tune.grid < 0 &&
          vnec_hip_strided_supported(tune.variant, tune.ipt, tune.encoding, plan_block(&tune), 1)))
        return vnec_fail("gathers issued together (lanes %d): variant C, FP64 or FP32, one right-hand side, "
                         "ipt 3, 5 or 7, per-tile carries (grid -1 or -2)", VNEC_LANES_GATHER);
    if (tune.variant == VNEC_VAR_G) {
        if (tune.nrhs != 1) return vnec_fail("variant G takes one right-hand side");
        tune.grid = -1;
    } else if (tune.variant == VNEC_VAR_C) {
        if (!valid_ipt(tune.ipt)) return vnec_fail("ipt %d: use an odd value in [3, 15]", tune.ipt);
        if (tune.encoding != VNEC_ENC_FP64 && tune.encoding != VNEC_ENC_FP32)
            return vnec_fail("variant C stores FP64 values (or FP32 for FP32 matrices)");
        if (tune.nrhs != 1) return vnec_fail("variant C takes one right-hand side");
    } else if (tune.variant == VNEC_VAR_H) {
        if (!vnec_hip_hot_supported(tune.ipt, tune.lanes))
            return vnec_fail("variant H: ipt 5 or 7 and lanes (hot entries) 4096 or 8192");
        if (tune.encoding != VNEC_ENC_FP64) return vnec_fail("variant H stores FP64 values only");
        if (tune.nrhs != 1) return vnec_fail("variant H takes one right-hand side");
        if (tune.grid < 0) tune.grid = 0; /* persistent only: the hot set loads once per workgroup */
    } else if (tune.variant == VNEC_VAR_D) {
        if (!vnec_hip_d_supported(tune.ipt, tune.lanes))
            return vnec_fail("variant D: ipt in {5, 7, 9} and lanes in {4, 8, 16}");
        if (tune.encoding != VNEC_ENC_FP64 && tune.encoding != VNEC_ENC_FP32)
            return vnec_fail("variant D stores FP64 values (or FP32 for FP32 matrices)");
        if (tune.nrhs != 1) return vnec_fail("variant D takes one right-hand side");
        tune.grid = -1;
    } else if (tune.variant == VNEC_VAR_S) {
        if (tune.ipt < 1 || tune.ipt > 7) return vnec_fail("variant S: ipt (tile = 2048 ipt items) in [1, 7]");
        if (tune.nrhs != 1) return vnec_fail("variant S takes one right-hand side");
    } else {
        if (!valid_ipt(tune.ipt)) return vnec_fail("ipt %d: use an odd value in [3, 15]", tune.ipt);
        if (!vnec_hip_merge_supported(tune.variant, tune.ipt, tune.encoding, tune.nrhs))
            return vnec_fail("no kernel for variant %d, ipt %d, encoding %d, nrhs %d", tune.variant,
                             tune.ipt, tune.encoding, tune.nrhs);
    }
    if (spb_matrix_upload(a) != SpB_SUCCESS) return SpB_FAILURE;

    struct vnec_plan_s *p = (struct vnec_plan_s *)calloc(1, sizeof *p);
    if (!p) return vnec_fail("out of memory");
    p->magic = SPB_MAGIC_PLAN;
    p->tune = tune;
    p->nrows = (int64_t)a->base.row;
    p->ncols = (int64_t)a->base.col;
    p->nnz = (int64_t)a->nnz;
    p->ptr = a->d_ptr;
    p->col = a->d_col;
    p->val = a->d_val;
    p->A = a;
    int64_t M = p->nrows + p->nnz;
    int64_t T = (int64_t)plan_threads(&tune) * tune.ipt;
    p->ntiles = M > 0 ? (M + T - 1) / T : 0;
    hipStream_t s = vnec_stream();
    SpB_Info rc = SpB_SUCCESS;
    if (tune.variant == VNEC_VAR_S) {
        p->ntiles = 0;
        p->tune.grid = -1;
        rc = p->nrows > 0 ? build_slices(p, a) : SpB_SUCCESS;
    } else if (tune.variant == VNEC_VAR_G) {
        rc = vnec_pattern_build(p, a);
        /* work units: pattern slices plus the sub-plan's tiles */
        p->ntiles = p->g_nslices + (p->g_subplan ? p->g_subplan->ntiles : 0);
    } else if (p->ntiles > 0) {
        p->coords = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * 2 * (size_t)(p->ntiles + 1), &p->dev_bytes);
        if (tune.nrhs == 1 && tune.grid >= 0 && tune.variant != VNEC_VAR_D) {
            int nwg = tune.grid;
            if (nwg == 0 && tune.variant == VNEC_VAR_H) {
                if (vnec_hip_ok(vnec_hip_hot_grid(tune.ipt, tune.lanes, &nwg), "occupancy")) rc = SpB_FAILURE;
            } else if (nwg == 0 && vnec_hip_ok(vnec_hip_stream_grid(tune.variant, tune.ipt, tune.encoding,
                                                                    plan_block(&tune), &nwg),
                                               "occupancy")) {
                rc = SpB_FAILURE;
            }
            nwg = nwg / 8 * 8;
            if (nwg < 8) nwg = 8;
            p->tune.grid = nwg;
            if (tune.variant == VNEC_VAR_H) nwg *= 4; /* four tile pipelines per workgroup */
            p->nwg = nwg;
            p->wg_row = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)nwg, &p->dev_bytes);
            p->wg_carry = (double *)vnec_malloc_dev(sizeof(double) * (size_t)nwg, &p->dev_bytes);
            if (!p->wg_row || !p->wg_carry) rc = SpB_FAILURE;
        } else {
            p->tune.grid = -1;
            p->carry = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(p->ntiles * tune.nrhs), &p->dev_bytes);
            if (!p->carry) rc = SpB_FAILURE;
            /* per-tile carries from a persistent launch when asked for, or
             * when each resident workgroup would see at least
             * STRIDED_MIN_TILES_PER_WG tiles */
            int snwg = 0;
            if (rc == SpB_SUCCESS && tune.nrhs == 1 &&
                (tune.variant == VNEC_VAR_L || tune.variant == VNEC_VAR_C) &&
                vnec_hip_strided_grid(tune.variant, tune.ipt, tune.encoding, plan_block(&tune), plan_gather(&tune),
                                      &snwg) == 0 &&
                (tune.grid == -2 || plan_block(&tune) || plan_gather(&tune) ||
                 p->ntiles >= STRIDED_MIN_TILES_PER_WG * (int64_t)snwg)) {
                p->snwg = snwg;
                p->tune.grid = -2;
            }
        }
        if (!p->coords) rc = SpB_FAILURE;
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(vnec_hip_partition(a->d_ptr, p->nrows, p->nnz, p->ntiles, 0, p->coords, s),
                             "partition");
        if (rc == SpB_SUCCESS && plan_relabelled(&p->tune)) rc = build_relabel(p, a);
        if (rc == SpB_SUCCESS && tune.variant == VNEC_VAR_D) rc = build_d(p, a);
        if (rc == SpB_SUCCESS && tune.variant == VNEC_VAR_H) rc = build_hot(p, a);
        if (rc == SpB_SUCCESS && tune.variant == VNEC_VAR_C) rc = build_csort(p, a);
        if (rc == SpB_SUCCESS && tune.variant == VNEC_VAR_E) {
            int32_t max_runs = 0;
            rc = build_windows(p, a, &max_runs, NULL);
            if (rc == SpB_SUCCESS && p->nwg > 0 && max_runs > VNEC_E_STREAM_MAX_RUNS) {
                /* too many runs in a tile for the pipelined kernel */
                vnec_free_dev(p->wg_row);
                vnec_free_dev(p->wg_carry);
                p->wg_row = NULL;
                p->wg_carry = NULL;
                p->nwg = 0;
                p->tune.grid = -1;
                p->carry = (double *)vnec_malloc_dev(sizeof(double) * (size_t)p->ntiles, &p->dev_bytes);
                if (!p->carry) rc = SpB_FAILURE;
            }
        }
        if (rc == SpB_SUCCESS && tune.encoding != VNEC_ENC_FP64 && p->nnz > 0) {
            p->hi = (float *)vnec_malloc_dev(sizeof(float) * (size_t)p->nnz, &p->dev_bytes);
            if (tune.encoding == VNEC_ENC_HILO)
                p->lo = (float *)vnec_malloc_dev(sizeof(float) * (size_t)p->nnz, &p->dev_bytes);
            if (!p->hi || (tune.encoding == VNEC_ENC_HILO && !p->lo)) rc = SpB_FAILURE;
            /* column-sorted tiles (FP32 matrices): the sorted values, whose
             * FP64 copy is then dropped (FP32 values widen and narrow exactly) */
            const double *src = tune.variant == VNEC_VAR_C ? p->cs_val : a->d_val;
            if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_split_hilo(src, p->nnz, p->hi, p->lo, s), "split");
            if (rc == SpB_SUCCESS && tune.variant == VNEC_VAR_C) {
                rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
                vnec_free_dev(p->cs_val);
                p->cs_val = NULL;
                p->dev_bytes -= (int64_t)sizeof(double) * p->nnz;
            }
        }
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "plan sync");
    }
    if (rc != SpB_SUCCESS) {
        plan_release(p);
        free(p);
        return SpB_FAILURE;
    }
    *plan = p;
    return SpB_SUCCESS;
}

/* Bytes of one value slot in the plan's encoding. */
static int64_t value_bytes(int enc) { return enc == VNEC_ENC_HI || enc == VNEC_ENC_FP32 ? 4 : 8; }

/* Bytes the main kernel of plan p must move per launch: every array of its
 * layout read once, x read once (every column), y written once.  Work
 * indices that are not per nonzero or per row (tile coordinates, window
 * counts, workgroup carries) are counted too; they are below 0.1 %. */
static int64_t main_kernel_bytes(const struct vnec_plan_s *p)
{
    const int enc = p->tune.encoding, k = p->tune.nrhs > 0 ? p->tune.nrhs : 1;
    const int64_t vb = value_bytes(enc), xb = enc == VNEC_ENC_FP32 ? 4 : 8;
    const int64_t nr = p->nrows, nc = p->ncols, nnz = p->nnz;
    const int64_t vecs = xb * k * (nc + nr), coords = 8 * (p->ntiles + 1);
    const int64_t windows = 8 * p->nruns + 4 * (p->ntiles + 1) + 4 * p->ntiles; /* runs, run_off, nec */
    switch (p->tune.variant) {
    case VNEC_VAR_G:
        return p->g_nslices * 64 * (int64_t)p->g_L * vb + 4 * nr + vecs;
    case VNEC_VAR_S:
        return p->s_padded * (vb + (p->win ? 2 : 4)) + 8 * (p->nslices + 1) + coords +
               (p->win ? windows : 0) + vecs;
    case VNEC_VAR_D:
        return ((p->groups + 63) / 64) * 64 * (int64_t)p->tune.lanes * vb + 4 * p->groups +
               4 * (nr + 1) + coords + windows + vecs;
    case VNEC_VAR_E:
        return nnz * (vb + 2) + 4 * (nr + 1) + coords + windows + vecs;
    case VNEC_VAR_C:
        return nnz * (vb + 4 + (p->cs_pos ? 2 : 0)) + 4 * (nr + 1) + coords +
               (p->cs_base ? 4 * p->ntiles : 0) + vecs;
    default: /* L, H: CSR with int32 columns */
        return nnz * (vb + 4) + 4 * (nr + 1) + coords + vecs;
    }
}

SpB_Info vnec_plan_bytes_ext(vnec_plan plan, int64_t *main_bytes, int64_t *spmv_bytes)
{
    if (!plan || plan->magic != SPB_MAGIC_PLAN) return vnec_fail("vnec_plan_bytes_ext: not a plan");
    int64_t mb = plan->ntiles > 0 ? main_kernel_bytes(plan) : 0, tot = mb;
    if (plan->tune.variant == VNEC_VAR_G && plan->g_nirr && plan->g_subplan) {
        /* the irregular rows: their plan (x already counted), its rows
         * written to scratch, read back and scattered with their indices */
        const int64_t xb = plan->tune.encoding == VNEC_ENC_FP32 ? 4 : 8, n = plan->g_nirr;
        int64_t sub = 0;
        vnec_plan_bytes_ext(plan->g_subplan, NULL, &sub);
        tot += sub - xb * plan->g_subplan->ncols + (2 * xb + 4) * n;
    }
    if (plan->cperm && plan->ntiles > 0) /* the x permutation: order and x read, permuted x written */
        tot += (4 + 8 + 8) * plan->ncols;
    if (main_bytes) *main_bytes = mb;
    if (spmv_bytes) *spmv_bytes = tot;
    return SpB_SUCCESS;
}

SpB_Info vnec_plan_info_ext(vnec_plan plan, vnec_tuning *t, int64_t *ntiles, int64_t *dev_bytes)
{
    if (!plan || plan->magic != SPB_MAGIC_PLAN) return vnec_fail("vnec_plan_info_ext: not a plan");
    if (t) *t = plan->tune;
    if (ntiles) *ntiles = plan->ntiles;
    if (dev_bytes) *dev_bytes = plan->dev_bytes;
    return SpB_SUCCESS;
}

/* Variant C: each tile's nonzeros re-stored in column order. */
static SpB_Info build_csort(struct vnec_plan_s *p, spb_csr *a)
{
    const size_t n = (size_t)(p->nnz > 0 ? p->nnz : 1);
    /* offsets within a tile stay below 256 ipt; when a column and an offset
     * fit one 32-bit word together, the offset rides in the column's high
     * bits (12 bytes per nonzero, as CSR) */
    int cb = 1, ob = 1;
    while (((int64_t)1 << cb) < p->ncols) cb++;
    while ((1 << ob) < plan_threads(&p->tune) * p->tune.ipt) ob++;
    p->cs_shift = cb + ob <= 32 ? cb : 0;
    if (!p->cs_shift && p->nnz > 0) {
        /* too many columns to pack: columns relative to each tile's smallest
         * one still fit when every tile spans few enough of them (a batch of
         * block-diagonal members: a tile lies within one or two members) */
        p->cs_base = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)p->ntiles, &p->dev_bytes);
        int32_t span = 0;
        if (!p->cs_base ||
            vnec_hip_ok(vnec_hip_tile_colspan(p->coords, p->ntiles, p->col, p->cs_base, &span, vnec_stream()),
                        "tile column span"))
            return SpB_FAILURE;
        int rb = 1;
        while (((int64_t)1 << rb) < (int64_t)span) rb++;
        if (rb + ob <= 32) {
            p->cs_shift = rb;
        } else {
            vnec_free_dev(p->cs_base);
            p->cs_base = NULL;
            p->dev_bytes -= (int64_t)sizeof(int32_t) * p->ntiles;
        }
    }
    p->cs_val = (double *)vnec_malloc_dev(sizeof(double) * n, &p->dev_bytes);
    p->cs_col = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * n, &p->dev_bytes);
    if (!p->cs_shift) p->cs_pos = (uint16_t *)vnec_malloc_dev(sizeof(uint16_t) * n, &p->dev_bytes);
    uint64_t *keys = (uint64_t *)vnec_malloc_dev(sizeof(uint64_t) * n, NULL);
    int32_t *idx = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * n, NULL);
    SpB_Info rc = (p->cs_val && p->cs_col && (p->cs_shift || p->cs_pos) && keys && idx) ? SpB_SUCCESS
                                                                                          : SpB_FAILURE;
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_csort(p->coords, p->ntiles, p->col, a->d_val, p->nnz, p->ncols, p->cs_shift,
                                        p->cs_base, keys, idx, p->cs_val, p->cs_col, p->cs_pos, vnec_stream()),
                         "column sort");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(vnec_stream()), "sync");
    vnec_free_dev(keys);
    vnec_free_dev(idx);
    return rc;
}

static void fill_args(const struct vnec_plan_s *p, const void *x_dev, void *y_dev,
                      vnec_merge_args *out)
{
    vnec_merge_args m;
    memset(&m, 0, sizeof m);
    m.ptr = p->ptr;
    m.col = p->col;
    m.lcol = p->lcol;
    m.val = p->val;
    m.hi = p->hi;
    m.lo = p->lo;
    m.coords = p->coords;
    m.ntiles = p->ntiles;
    m.run_off = p->run_off;
    m.run_start = p->run_start;
    m.run_wpos = p->run_wpos;
    m.nec = p->nec;
    if (p->tune.encoding == VNEC_ENC_FP32) {
        m.x32 = (const float *)x_dev;
        m.y32 = (float *)y_dev;
        m.carry32 = (float *)p->carry; /* ntiles doubles hold ntiles floats */
    } else {
        m.x = (const double *)x_dev;
        m.y = (double *)y_dev;
        m.carry = p->carry;
    }
    m.nrows = p->nrows;
    m.ipt = p->tune.ipt;
    m.variant = p->tune.variant;
    m.encoding = p->tune.encoding;
    m.nrhs = p->tune.nrhs;
    if (p->tune.variant == VNEC_VAR_C) {
        m.val = p->cs_val;
        m.col = p->cs_col;
        m.spos = p->cs_pos;
        m.cpos_shift = p->cs_shift;
        m.cbase = p->cs_base;
        m.block = plan_block(&p->tune);
        m.gather = plan_gather(&p->tune);
    }
    if (p->tune.variant == VNEC_VAR_H) {
        m.col = p->hcol;
        m.xhot = p->xhot;
        m.nhot = p->nhot;
    }
    *out = m;
}

static SpB_Info launch_slice(struct vnec_plan_s *p, const void *x_dev, void *y_dev, hipStream_t s)
{
    vnec_slice_args a;
    memset(&a, 0, sizeof a);
    a.ptr = p->ptr;
    a.slice_len = p->slice_len;
    a.slice_ptr = p->slice_ptr;
    a.sv = p->sv;
    a.shi = p->shi;
    a.slo = p->slo;
    a.scol = p->scol;
    a.slcol = p->slcol;
    a.tile_slices = p->tile_slices;
    a.ntiles = p->ntiles;
    a.run_off = p->run_off;
    a.run_start = p->run_start;
    a.run_wpos = p->run_wpos;
    a.nec = p->nec;
    a.max_nec = p->max_nec;
    a.win = p->win;
    if (p->tune.encoding == VNEC_ENC_FP32) {
        a.x32 = (const float *)x_dev;
        a.y32 = (float *)y_dev;
    } else {
        a.x = (const double *)x_dev;
        a.y = (double *)y_dev;
    }
    a.nrows = p->nrows;
    a.encoding = p->tune.encoding;
    return vnec_hip_ok(vnec_hip_spmv_slice(&a, s), "slice kernel");
}

/* Launch the SpMV kernel only (its carries stay in the plan's scratch). */
static SpB_Info launch_d(struct vnec_plan_s *p, const void *x_dev, void *y_dev, hipStream_t s)
{
    vnec_d_args a;
    memset(&a, 0, sizeof a);
    a.ptr = p->ptr;
    a.vrp = p->vrp;
    a.col_start = p->col_start;
    a.vals = p->dvals;
    a.val = p->val;
    a.lcol = p->lcol;
    a.coords = p->coords;
    a.ntiles = p->ntiles;
    a.run_off = p->run_off;
    a.run_start = p->run_start;
    a.run_wpos = p->run_wpos;
    a.nec = p->nec;
    if (p->tune.encoding == VNEC_ENC_FP32) {
        a.fp32 = 1;
        a.vals32 = p->dvals32;
        a.val32 = p->hi;
        a.x32 = (const float *)x_dev;
        a.y32 = (float *)y_dev;
        a.carry32 = (float *)p->carry;
    } else {
        a.x = (const double *)x_dev;
        a.y = (double *)y_dev;
        a.carry = p->carry;
    }
    a.nrows = p->nrows;
    a.ipt = p->tune.ipt;
    a.lanes = p->tune.lanes;
    return vnec_hip_ok(vnec_hip_spmv_d(&a, s), "d kernel");
}

static SpB_Info launch_main(struct vnec_plan_s *p, const vnec_merge_args *m, const void *x_dev,
                            void *y_dev, hipStream_t s)
{
    if (p->tune.variant == VNEC_VAR_G) {
        vnec_pattern_args g;
        g.masks = p->g_masks;
        g.gv = p->g_vals;
        g.gh = p->g_hi;
        g.gl = p->g_lo;
        g.encoding = p->tune.encoding;
        g.pat = p->g_pat;
        g.L = p->g_L;
        g.nrows = p->nrows;
        g.nslices = p->g_nslices;
        g.x = (const double *)x_dev;
        g.y = (double *)y_dev;
        g.ncols = p->ncols;
        if (p->tune.encoding == VNEC_ENC_FP32) {
            g.x = NULL;
            g.y = NULL;
            g.x32 = (const float *)x_dev;
            g.y32 = (float *)y_dev;
        } else {
            g.x32 = NULL;
            g.y32 = NULL;
        }
        vnec_pattern_window(p->g_pat_h, p->g_L, &g);
        return vnec_hip_ok(vnec_hip_spmv_pattern(&g, s), "pattern kernel");
    }
    if (p->tune.variant == VNEC_VAR_S) return launch_slice(p, x_dev, y_dev, s);
    if (p->tune.variant == VNEC_VAR_D) return launch_d(p, x_dev, y_dev, s);
    if (p->tune.variant == VNEC_VAR_H) {
        if (vnec_hip_ok(vnec_hip_gather_hot(m->x, p->hot_order, p->nhot, p->xhot, s), "hot gather"))
            return SpB_FAILURE;
        return vnec_hip_ok(vnec_hip_spmv_hot(m, p->tune.lanes, p->tune.grid, p->wg_row, p->wg_carry, s),
                           "spmv kernel");
    }
    if (p->nwg > 0)
        return vnec_hip_ok(vnec_hip_spmv_stream(m, p->nwg, p->wg_row, p->wg_carry, s), "spmv kernel");
    if (p->snwg > 0) return vnec_hip_ok(vnec_hip_spmv_strided(m, p->snwg, s), "spmv kernel");
    return vnec_hip_ok(vnec_hip_spmv_merge(m, s), "spmv kernel");
}

static SpB_Info launch_fixup(struct vnec_plan_s *p, const void *x_dev, void *y_dev, hipStream_t s)
{
    if (p->tune.variant == VNEC_VAR_S) return SpB_SUCCESS; /* row-aligned tiles: no carries */
    if (p->tune.variant == VNEC_VAR_G) { /* the irregular rows */
        if (!p->g_nirr) return SpB_SUCCESS;
        if (vnec_plan_run(p->g_subplan, x_dev, p->g_ysub) != SpB_SUCCESS) return SpB_FAILURE;
        if (p->tune.encoding == VNEC_ENC_FP32)
            return vnec_hip_ok(vnec_hip_scatter_rows_f32((const float *)p->g_ysub, p->g_irr, p->g_nirr,
                                                         (float *)y_dev, s),
                               "scatter rows");
        return vnec_hip_ok(vnec_hip_scatter_rows(p->g_ysub, p->g_irr, p->g_nirr, (double *)y_dev, s),
                           "scatter rows");
    }
    if (p->tune.encoding == VNEC_ENC_FP32)
        return vnec_hip_ok(vnec_hip_carry_fixup_f32(p->coords, (const float *)p->carry, p->ntiles,
                                                    (float *)y_dev, p->nrows, s), "carry fix-up");
    if (p->nwg > 0)
        return vnec_hip_ok(vnec_hip_wg_fixup(p->wg_row, p->wg_carry, p->nwg, y_dev, p->nrows, s),
                           "carry fix-up");
    return vnec_hip_ok(vnec_hip_carry_fixup(p->coords, p->carry, p->ntiles, y_dev, p->nrows,
                                            p->tune.nrhs, s), "carry fix-up");
}

/* Enqueue one SpMV (kernel + fix-up) on the library stream. */
SpB_Info vnec_plan_run(struct vnec_plan_s *p, const void *x_dev, void *y_dev)
{
    hipStream_t s = vnec_stream();
    if (p->ntiles == 0) {
        if (p->tune.encoding == VNEC_ENC_FP32)
            return vnec_hip_ok(vnec_hip_zero_f32((float *)y_dev, p->nrows, s), "zero y");
        return vnec_hip_ok(vnec_hip_zero_f64((double *)y_dev, p->nrows * p->tune.nrhs, s), "zero y");
    }
    if (p->cperm) { /* x in the relabelled order */
        if (vnec_hip_ok(vnec_hip_gather_hot((const double *)x_dev, p->cperm, p->ncols, p->xperm, s), "x permute"))
            return SpB_FAILURE;
        x_dev = p->xperm;
    }
    vnec_merge_args m;
    fill_args(p, x_dev, y_dev, &m);
    if (launch_main(p, &m, x_dev, y_dev, s) != SpB_SUCCESS) return SpB_FAILURE;
    return launch_fixup(p, x_dev, y_dev, s);
}

SpB_Info vnec_spmv_dev_ext(vnec_plan plan, const void *x_dev, void *y_dev)
{
    if (!plan || plan->magic != SPB_MAGIC_PLAN) return vnec_fail("vnec_spmv_dev_ext: not a plan");
    return vnec_plan_run(plan, x_dev, y_dev);
}

/* One SpMV of p on (x_dev, y_dev) replayed from a HIP graph: captured on
 * the library stream the first time the pair is seen (the plan keeps
 * VNEC_PLAN_GRAPHS pairs, the least recently used one is recaptured), then
 * launched.  The launch sequence records no memset node. */
static SpB_Info plan_run_graph(struct vnec_plan_s *p, const void *x_dev, void *y_dev)
{
    hipStream_t s = vnec_stream();
    int slot = -1, lru = 0;
    for (int i = 0; i < VNEC_PLAN_GRAPHS; i++) {
        if (p->sgraph[i].exec && p->sgraph[i].x == x_dev && p->sgraph[i].y == y_dev) slot = i;
        if (p->sgraph[i].used < p->sgraph[lru].used) lru = i;
    }
    if (slot < 0) {
        slot = lru;
        if (p->sgraph[slot].exec) {
            hipGraphExecDestroy(p->sgraph[slot].exec);
            p->sgraph[slot].exec = NULL;
        }
        hipGraph_t g = NULL;
        SpB_Info rc = vnec_hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture");
        if (rc != SpB_SUCCESS) return rc;
        rc = vnec_plan_run(p, x_dev, y_dev);
        SpB_Info rc2 = vnec_hip_ok(hipStreamEndCapture(s, &g), "end capture");
        if (rc == SpB_SUCCESS) rc = rc2;
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(hipGraphInstantiate(&p->sgraph[slot].exec, g, NULL, NULL, 0), "instantiate");
        if (g) hipGraphDestroy(g);
        if (rc != SpB_SUCCESS) {
            p->sgraph[slot].exec = NULL;
            return rc;
        }
        p->sgraph[slot].x = x_dev;
        p->sgraph[slot].y = y_dev;
    }
    p->sgraph[slot].used = ++p->sgraph_clock;
    return vnec_hip_ok(hipGraphLaunch(p->sgraph[slot].exec, s), "graph launch");
}

SpB_Info vnec_spmv_graph_ext(vnec_plan plan, const void *x_dev, void *y_dev)
{
    if (!plan || plan->magic != SPB_MAGIC_PLAN) return vnec_fail("vnec_spmv_graph_ext: not a plan");
    if (!x_dev || !y_dev) return vnec_fail("vnec_spmv_graph_ext: NULL vector");
    return plan_run_graph(plan, x_dev, y_dev);
}

SpB_Info vnec_spmm_colmajor_dev_ext(vnec_plan plan, const void *X_dev, void *Y_dev)
{
    if (!plan || plan->magic != SPB_MAGIC_PLAN) return vnec_fail("vnec_spmm_colmajor_dev_ext: not a plan");
    const int k = plan->tune.nrhs;
    if (k <= 1) return vnec_plan_run(plan, X_dev, Y_dev);
    if (!plan->xt) {
        plan->xt = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(plan->ncols > 0 ? plan->ncols * k : 1),
                                             &plan->dev_bytes);
        plan->yt = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(plan->nrows > 0 ? plan->nrows * k : 1),
                                             &plan->dev_bytes);
        if (!plan->xt || !plan->yt) return SpB_FAILURE;
    }
    hipStream_t s = vnec_stream();
    if (vnec_hip_ok(vnec_hip_rhs_layout((const double *)X_dev, plan->ncols, k, 1, plan->xt, s), "interleave X"))
        return SpB_FAILURE;
    if (vnec_plan_run(plan, plan->xt, plan->yt) != SpB_SUCCESS) return SpB_FAILURE;
    return vnec_hip_ok(vnec_hip_rhs_layout(plan->yt, plan->nrows, k, 0, (double *)Y_dev, s), "de-interleave Y");
}

SpB_Info vnec_spmv_ext(vnec_plan plan, SpB_Vector y, const SpB_Vector x)
{
    if (!plan || plan->magic != SPB_MAGIC_PLAN) return vnec_fail("vnec_spmv_ext: not a plan");
    spb_dense *xv = spb_as_dense(x), *yv = spb_as_dense(y);
    if (!xv || !yv) return SpB_FAILURE;
    int k = plan->tune.nrhs;
    SpB_Type vt = plan->tune.encoding == VNEC_ENC_FP32 ? SpB_FP32 : SpB_FP64;
    if (xv->base.element_type != vt || yv->base.element_type != vt)
        return vnec_fail("vnec_spmv_ext: the plan computes in %s; x and y must be %s vectors",
                         vt == SpB_FP32 ? "FP32" : "FP64", vt == SpB_FP32 ? "SpB_FP32" : "SpB_FP64");
    if ((int64_t)xv->base.n != plan->ncols * k || (int64_t)yv->base.n != plan->nrows * k)
        return vnec_fail("vnec_spmv_ext: x has %lu entries (need %ld), y %lu (need %ld)",
                         xv->base.n, (long)(plan->ncols * k), yv->base.n, (long)(plan->nrows * k));
    if (spb_vector_to_device(xv) != SpB_SUCCESS) return SpB_FAILURE;
    if (spb_vector_device_alloc(yv) != SpB_SUCCESS) return SpB_FAILURE;
    /* direct launches: replaying a captured graph per call measured slower
     * (one hipGraphLaunch costs more than the plan's one to three kernel
     * launches), so vnec_spmv_graph_ext stays an explicit choice */
    if (vnec_plan_run(plan, xv->d_values, yv->d_values) != SpB_SUCCESS) return SpB_FAILURE;
    yv->dev_stale = 0;
    yv->host_stale = 1;
    if (!yv->resident) return spb_vector_from_device(yv);
    return vnec_hip_ok(hipStreamSynchronize(vnec_stream()), "spmv sync");
}

/* ---------------------------------------------------------------- timing */
static void *g_flush = NULL;
static size_t g_flush_bytes = 0;

SpB_Info vnec_time_dev_ext(vnec_plan plan, const void *x_dev, void *y_dev, int iters,
                           int use_graph, size_t flush_bytes, double *ms)
{
    if (!plan || plan->magic != SPB_MAGIC_PLAN || iters < 1 || !ms)
        return vnec_fail("vnec_time_dev_ext: bad arguments");
    hipStream_t s = vnec_stream();
    hipEvent_t e0, e1;
    if (vnec_hip_ok(hipEventCreate(&e0), "event") || vnec_hip_ok(hipEventCreate(&e1), "event"))
        return SpB_FAILURE;
    SpB_Info rc = SpB_SUCCESS;
    flush_bytes = flush_bytes / 8 * 8;
    if (flush_bytes > 0 && flush_bytes != g_flush_bytes) {
        vnec_free_dev(g_flush);
        g_flush = vnec_malloc_dev(flush_bytes + 16384, NULL); /* + the probe's output words */
        g_flush_bytes = g_flush ? flush_bytes : 0;
        if (!g_flush) rc = SpB_FAILURE;
        else rc = vnec_hip_ok(hipMemsetAsync(g_flush, 1, flush_bytes + 16384, s), "memset");
    }
    /* warm-up */
    if (rc == SpB_SUCCESS) rc = vnec_plan_run(plan, x_dev, y_dev);
    if (rc == SpB_SUCCESS && use_graph && !flush_bytes &&
        (!plan->gexec || plan->g_x != x_dev || plan->g_y != y_dev)) {
        if (plan->gexec) {
            hipGraphExecDestroy(plan->gexec);
            plan->gexec = NULL;
        }
        int reps = iters < 20 ? iters : 20;
        hipGraph_t g;
        rc = vnec_hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture");
        for (int r = 0; rc == SpB_SUCCESS && r < reps; r++) rc = vnec_plan_run(plan, x_dev, y_dev);
        SpB_Info rc2 = vnec_hip_ok(hipStreamEndCapture(s, &g), "end capture");
        if (rc == SpB_SUCCESS) rc = rc2;
        if (rc == SpB_SUCCESS) {
            rc = vnec_hip_ok(hipGraphInstantiate(&plan->gexec, g, NULL, NULL, 0), "instantiate");
            hipGraphDestroy(g);
        }
        if (rc == SpB_SUCCESS) {
            plan->g_x = x_dev;
            plan->g_y = y_dev;
            plan->g_reps = reps;
            rc = vnec_hip_ok(hipGraphLaunch(plan->gexec, s), "graph warm-up");
        }
    }
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "warm-up sync");
    double total_ms = 0.0;
    int done = 0;
    if (rc == SpB_SUCCESS && flush_bytes) {
        /* iters x (flush, SpMV) back to back, less iters flushes alone: the
         * flush reads a buffer larger than L2 and the Infinity Cache with
         * cache-allocating loads, so every SpMV starts from HBM, and one
         * pair of events spans each sequence */
        void *out = (char *)g_flush + flush_bytes;
        float f_both = 0.f, f_flush = 0.f;
        for (int pass = 0; pass < 2 && rc == SpB_SUCCESS; pass++) {
            rc = vnec_hip_ok(vnec_hip_flush_read(g_flush, flush_bytes, out, s), "flush");
            if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventRecord(e0, s), "record");
            for (int i = 0; rc == SpB_SUCCESS && i < iters; i++) {
                rc = vnec_hip_ok(vnec_hip_flush_read(g_flush, flush_bytes, out, s), "flush");
                if (rc == SpB_SUCCESS && pass == 0) rc = vnec_plan_run(plan, x_dev, y_dev);
            }
            if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventRecord(e1, s), "record");
            if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventSynchronize(e1), "sync");
            if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventElapsedTime(pass ? &f_flush : &f_both, e0, e1), "elapsed");
        }
        total_ms = (double)f_both - (double)f_flush;
        done = iters;
    } else if (rc == SpB_SUCCESS) {
        rc = vnec_hip_ok(hipEventRecord(e0, s), "record");
        if (use_graph && plan->gexec) {
            for (done = 0; rc == SpB_SUCCESS && done < iters; done += plan->g_reps)
                rc = vnec_hip_ok(hipGraphLaunch(plan->gexec, s), "graph launch");
        } else {
            for (done = 0; rc == SpB_SUCCESS && done < iters; done++)
                rc = vnec_plan_run(plan, x_dev, y_dev);
        }
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventRecord(e1, s), "record");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventSynchronize(e1), "sync");
        float f = 0.f;
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventElapsedTime(&f, e0, e1), "elapsed");
        total_ms = f;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc != SpB_SUCCESS) return rc;
    *ms = total_ms / (done > 0 ? done : 1);
    return SpB_SUCCESS;
}

/* Mean duration of the SpMV kernel and of the fix-up kernel on the library
 * stream: iters launches of the SpMV kernel back to back between one pair of
 * events, then iters launches of the fix-up between the next pair (a pair of
 * events per launch would add the events' own few microseconds to each).  The
 * kernels keep no state between launches, so repeating one alone is safe; the
 * repeated fix-ups leave y wrong, and a final SpMV puts it right. */
SpB_Info vnec_time_kernels_ext(vnec_plan plan, const void *x_dev, void *y_dev, int iters,
                               double *ms_main, double *ms_fixup)
{
    if (!plan || plan->magic != SPB_MAGIC_PLAN || iters < 1)
        return vnec_fail("vnec_time_kernels_ext: bad arguments");
    if (plan->ntiles == 0) {
        if (ms_main) *ms_main = 0.0;
        if (ms_fixup) *ms_fixup = 0.0;
        return SpB_SUCCESS;
    }
    hipStream_t s = vnec_stream();
    hipEvent_t ev[3];
    for (int i = 0; i < 3; i++)
        if (vnec_hip_ok(hipEventCreate(&ev[i]), "event")) return SpB_FAILURE;
    vnec_merge_args m;
    /* a relabelled plan's kernels read the permuted x, which the SpMV just
     * before the timed launches leaves in the plan's scratch */
    fill_args(plan, plan->cperm ? plan->xperm : x_dev, y_dev, &m);
    SpB_Info rc = vnec_plan_run(plan, x_dev, y_dev);
    float a = 0.f, b = 0.f;
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventRecord(ev[0], s), "record");
    for (int i = 0; rc == SpB_SUCCESS && i < iters; i++)
        rc = launch_main(plan, &m, x_dev, y_dev, s);
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventRecord(ev[1], s), "record");
    for (int i = 0; rc == SpB_SUCCESS && i < iters; i++)
        rc = launch_fixup(plan, x_dev, y_dev, s);
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventRecord(ev[2], s), "record");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventSynchronize(ev[2]), "sync");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventElapsedTime(&a, ev[0], ev[1]), "elapsed");
    if (rc == SpB_SUCCESS && plan->tune.variant != VNEC_VAR_S) /* S has no fix-up kernel */
        rc = vnec_hip_ok(hipEventElapsedTime(&b, ev[1], ev[2]), "elapsed");
    if (rc == SpB_SUCCESS) rc = vnec_plan_run(plan, x_dev, y_dev);
    for (int i = 0; i < 3; i++) hipEventDestroy(ev[i]);
    if (rc != SpB_SUCCESS) return rc;
    if (ms_main) *ms_main = (double)a / iters;
    if (ms_fixup) *ms_fixup = (double)b / iters;
    return SpB_SUCCESS;
}

/* Slots of the 64-row slice layout per stored entry (1.0: no padding). */
double vnec_slice_padding(const spb_csr *a)
{
    int64_t nr = (int64_t)a->base.row, slots = 0;
    for (int64_t s = 0; s < nr; s += 64) {
        int64_t e = s + 64 < nr ? s + 64 : nr, m = 0;
        for (int64_t r = s; r < e; r++) {
            int64_t l = (int64_t)(a->ptr[r + 1] - a->ptr[r]);
            m = l > m ? l : m;
        }
        slots += 64 * m;
    }
    return a->nnz ? (double)slots / (double)a->nnz : 1.0;
}

/* Streaming probe over two buffers of `bytes`: a copy (read + write) or a
 * read alone; *gbps counts the bytes moved per launch over the mean launch
 * time of iters back-to-back launches. */
static SpB_Info probe_bandwidth(size_t bytes, int iters, int read_only, double *gbps, const char *fn)
{
    if (!gbps || iters < 1 || bytes < 16) return vnec_fail("%s: bad arguments", fn);
    if (!vnec_runtime_ready()) return SpB_FAILURE;
    bytes = bytes / 16 * 16;
    void *a = vnec_malloc_dev(bytes, NULL), *b = vnec_malloc_dev(bytes, NULL);
    hipStream_t s = vnec_stream();
    hipEvent_t e0 = NULL, e1 = NULL;
    SpB_Info rc = (a && b) ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemsetAsync(a, 1, bytes, s), "memset");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventCreate(&e0), "event");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventCreate(&e1), "event");
    for (int i = -1; rc == SpB_SUCCESS && i < iters; i++) { /* i = -1: warm-up */
        if (i == 0) rc = vnec_hip_ok(hipEventRecord(e0, s), "record");
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(read_only ? vnec_hip_stream_read(a, bytes, b, s) : vnec_hip_stream_copy(a, b, bytes, s),
                             "probe");
    }
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventRecord(e1, s), "record");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventSynchronize(e1), "sync");
    float ms = 0.f;
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    if (rc == SpB_SUCCESS) *gbps = (read_only ? 1.0 : 2.0) * (double)bytes * iters / (ms / 1e3) / 1e9;
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    vnec_free_dev(a);
    vnec_free_dev(b);
    return rc;
}

SpB_Info vnec_copy_bandwidth_ext(size_t bytes, int iters, double *gbps)
{
    return probe_bandwidth(bytes, iters, 0, gbps, "vnec_copy_bandwidth_ext");
}

SpB_Info vnec_read_bandwidth_ext(size_t bytes, int iters, double *gbps)
{
    return probe_bandwidth(bytes, iters, 1, gbps, "vnec_read_bandwidth_ext");
}

#define GATHER_PROBE_GRID 2048
SpB_Info vnec_gather_ceiling_ext(const SpB_Matrix A, const double *x_dev, int iters, double *stream_ms,
                                 double *gather_ms)
{
    spb_csr *a = spb_as_csr(A);
    if (!a || !x_dev || iters < 1 || !stream_ms || !gather_ms)
        return vnec_fail("vnec_gather_ceiling_ext: bad arguments");
    if (a->base.element_type != SpB_FP64) return vnec_fail("vnec_gather_ceiling_ext: FP64 matrices only");
    if (spb_matrix_upload(a) != SpB_SUCCESS) return SpB_FAILURE;
    hipStream_t s = vnec_stream();
    double *out = (double *)vnec_malloc_dev(sizeof(double) * GATHER_PROBE_GRID * 256, NULL);
    hipEvent_t e0 = NULL, e1 = NULL;
    SpB_Info rc = out ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventCreate(&e0), "event");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventCreate(&e1), "event");
    for (int gather = 0; gather < 2 && rc == SpB_SUCCESS; gather++) {
        rc = vnec_hip_ok(vnec_hip_gather_probe(a->d_val, a->d_col, (int64_t)a->nnz, x_dev, gather, out,
                                               GATHER_PROBE_GRID, s), "probe"); /* warm-up */
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventRecord(e0, s), "record");
        for (int i = 0; i < iters && rc == SpB_SUCCESS; i++)
            rc = vnec_hip_ok(vnec_hip_gather_probe(a->d_val, a->d_col, (int64_t)a->nnz, x_dev, gather, out,
                                                   GATHER_PROBE_GRID, s), "probe");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventRecord(e1, s), "record");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventSynchronize(e1), "sync");
        float ms = 0.f;
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        if (rc == SpB_SUCCESS) *(gather ? gather_ms : stream_ms) = ms / iters;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    vnec_free_dev(out);
    return rc;
}
