/*
 * spb_vnec.c — the reference's VNEC entry points (IRD_VNEC, Build_VNEC_*,
 * SpMV_VNEC_*_FP64 and their FP32 twins) and SpB_mxv, on the GPU.
 *
 * Each Build_VNEC_* call returns the reference's public plan struct with
 * its host fields filled at the reference's compiled parameters (P = 36
 * partitions, L = 8 lanes for FP64 and 16 for FP32,
 * src/operation/include/SpB_mxv.h:26-33), computed by the GPU builders,
 * and keeps behind it a GPU execution plan laid out for wave64 tiles.
 * SpMV_VNEC_* runs that GPU plan; y is valid on return.  The FP32 entry
 * points compute as the reference's FP32 kernels do: FP32 values, x, y and
 * sums.  The D plans' val_align holds G * L values (G = the group count,
 * 64-byte aligned, freed with free()), the reference's layout without its
 * unused tail: the reference allocates nnz * L entries and fills G * L.
 */
#include <stdlib.h>
#include <string.h>

#include "spb_internal.h"

#define REF_P 36     /* NUM_THREADS, SpB_mxv.h:28 */
#define REF_L 8      /* nLanes_f64,  SpB_mxv.h:30 */
#define REF_L32 16   /* nLanes_f32,  SpB_mxv.h:29 */
#define REF_IRD_THR 0.477   /* IRD_thr_fp64, SpB_mxv.h:32 */
#define REF_IRD_THR32 0.437 /* IRD_thr_fp32, SpB_mxv.h:31 */
#define REF_CROSS_MAX 32  /* longest row sent to the cross-row kernel, spmv_vnec_S_for_mxv.c:489 */

typedef enum { KIND_L = 1, KIND_D = 2, KIND_S = 3, KIND_D32 = 4, KIND_S32 = 5 } plan_kind;

/* Registry of public plans handed out, so SpB_mxv can tell them apart. */
typedef struct wrap {
    void *pub;
    plan_kind kind;
    struct vnec_plan_s *gpu;
    const spb_csr *A;
    struct wrap *next;
} wrap;
static wrap *g_wraps = NULL;

static wrap *find_wrap(const void *pub)
{
    for (wrap *w = g_wraps; w; w = w->next)
        if (w->pub == pub) return w;
    return NULL;
}

static void add_wrap(void *pub, plan_kind kind, struct vnec_plan_s *gpu, const spb_csr *A)
{
    wrap *w = (wrap *)calloc(1, sizeof *w);
    w->pub = pub;
    w->kind = kind;
    w->gpu = gpu;
    w->A = A;
    w->next = g_wraps;
    g_wraps = w;
}

static struct vnec_plan_s *drop_wrap(const void *pub)
{
    wrap **pp = &g_wraps;
    while (*pp) {
        if ((*pp)->pub == pub) {
            wrap *w = *pp;
            struct vnec_plan_s *g = w->gpu;
            *pp = w->next;
            free(w);
            return g;
        }
        pp = &(*pp)->next;
    }
    return NULL;
}

vnec_plan VNEC_plan_of_ext(const void *reference_plan)
{
    wrap *w = find_wrap(reference_plan);
    return w ? w->gpu : NULL;
}

/* ------------------------------------------------ structure exports */
static int64_t ipt_of(int64_t nrows, int64_t nnz, int64_t P)
{
    int64_t M = nrows + nnz;
    return (M + P - 1) / P;
}

SpB_Info vnec_partition_ext(const SpB_Matrix A, int64_t P, int snap, int32_t *coords)
{
    spb_csr *a = spb_as_csr(A);
    if (!a || !coords || P < 1) return vnec_fail("vnec_partition_ext: bad arguments");
    if (spb_matrix_upload(a) != SpB_SUCCESS) return SpB_FAILURE;
    hipStream_t s = vnec_stream();
    int32_t *d = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * 2 * (size_t)(P + 1), NULL);
    if (!d) return SpB_FAILURE;
    SpB_Info rc = vnec_hip_ok(vnec_hip_partition(a->d_ptr, (int64_t)a->base.row, (int64_t)a->nnz, P,
                                                 snap, d, s), "partition");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(coords, d, sizeof(int32_t) * 2 * (size_t)(P + 1),
                                        hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    vnec_free_dev(d);
    return rc;
}

/* Device ECR at the partitions in d_coords (device).  Writes device ecr
 * (nnz), win (<= nnz), win_off (P + 1) and host nec sums. */
static SpB_Info ecr_device(spb_csr *a, int64_t P, const int32_t *d_coords, int method,
                           int32_t *d_ecr, int32_t *d_win, int64_t *d_win_off)
{
    hipStream_t s = vnec_stream();
    int64_t nnz = (int64_t)a->nnz, ncols = (int64_t)a->base.col;
    int32_t mx = 0;
    if (vnec_hip_ok(vnec_hip_max_tile_nnz(d_coords, P, &mx, s), "max part nnz")) return SpB_FAILURE;
    int dense_ok = P * (ncols > 0 ? ncols : 1) < ((int64_t)1 << 31) - 1;
    int tile_ok = mx <= VNEC_ECR_TILE_MAX;
    if (method == 0) method = tile_ok ? 2 : 1;
    if (method == 1 && !dense_ok) return vnec_fail("dense ECR needs P * ncols < 2^31");
    if (method == 2 && !tile_ok) return vnec_fail("tile ECR needs partitions of <= %d nonzeros", VNEC_ECR_TILE_MAX);
    if (method == 1) {
        int64_t tot = P * ncols;
        int32_t *flags = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(tot ? tot : 1), NULL);
        int32_t *rank = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(tot ? tot : 1), NULL);
        SpB_Info rc = (flags && rank) ? SpB_SUCCESS : SpB_FAILURE;
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(vnec_hip_ecr_dense(a->d_col, nnz, ncols, d_coords, P, flags, rank, d_ecr,
                                                d_win, d_win_off, s), "dense ecr");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
        vnec_free_dev(flags);
        vnec_free_dev(rank);
        return rc;
    }
    /* tile form: uniq lists sit at each partition's first nonzero; compact */
    int32_t *uniq = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nnz ? nnz : 1), NULL);
    int32_t *nec = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(P + 1), NULL);
    int32_t *nruns = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(P + 1), NULL);
    SpB_Info rc = (uniq && nec && nruns) ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_ecr_tiles_pass1(a->d_col, d_coords, P, NULL, d_ecr, uniq, nec, nruns, s),
                         "tile ecr");
    int32_t *hnec = NULL, *hco = NULL;
    int64_t *hoff = NULL;
    if (rc == SpB_SUCCESS) {
        hnec = (int32_t *)malloc(sizeof(int32_t) * (size_t)P);
        hco = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(P + 1));
        hoff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(P + 1));
        rc = (hnec && hco && hoff) ? SpB_SUCCESS : vnec_fail("out of memory");
    }
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(hnec, nec, sizeof(int32_t) * (size_t)P, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(hco, d_coords, sizeof(int32_t) * 2 * (size_t)(P + 1),
                                        hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    if (rc == SpB_SUCCESS) {
        int64_t off = 0;
        for (int64_t t = 0; t < P && rc == SpB_SUCCESS; t++) {
            hoff[t] = off;
            if (hnec[t] > 0)
                rc = vnec_hip_ok(hipMemcpyAsync(d_win + off, uniq + hco[2 * t + 1],
                                                sizeof(int32_t) * (size_t)hnec[t],
                                                hipMemcpyDeviceToDevice, s), "d2d");
            off += hnec[t];
        }
        hoff[P] = off;
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(hipMemcpyAsync(d_win_off, hoff, sizeof(int64_t) * (size_t)(P + 1),
                                            hipMemcpyHostToDevice, s), "h2d");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    }
    free(hnec);
    free(hco);
    free(hoff);
    vnec_free_dev(uniq);
    vnec_free_dev(nec);
    vnec_free_dev(nruns);
    return rc;
}

SpB_Info vnec_ecr_ext(const SpB_Matrix A, int64_t P, const int32_t *coords, int method, int32_t *ecr,
                      int32_t *nec, int64_t *win_off, int32_t *win)
{
    spb_csr *a = spb_as_csr(A);
    if (!a || !coords || P < 1 || !ecr || !nec || !win_off || !win)
        return vnec_fail("vnec_ecr_ext: bad arguments");
    if (spb_matrix_upload(a) != SpB_SUCCESS) return SpB_FAILURE;
    hipStream_t s = vnec_stream();
    int64_t nnz = (int64_t)a->nnz;
    int32_t *dco = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * 2 * (size_t)(P + 1), NULL);
    int32_t *decr = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nnz ? nnz : 1), NULL);
    int32_t *dwin = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nnz ? nnz : 1), NULL);
    int64_t *doff = (int64_t *)vnec_malloc_dev(sizeof(int64_t) * (size_t)(P + 1), NULL);
    SpB_Info rc = (dco && decr && dwin && doff) ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(dco, coords, sizeof(int32_t) * 2 * (size_t)(P + 1),
                                        hipMemcpyHostToDevice, s), "h2d");
    if (rc == SpB_SUCCESS) rc = ecr_device(a, P, dco, method, decr, dwin, doff);
    if (rc == SpB_SUCCESS && nnz)
        rc = vnec_hip_ok(hipMemcpyAsync(ecr, decr, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(win_off, doff, sizeof(int64_t) * (size_t)(P + 1), hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    if (rc == SpB_SUCCESS && win_off[P] > 0)
        rc = vnec_hip_ok(hipMemcpyAsync(win, dwin, sizeof(int32_t) * (size_t)win_off[P], hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    if (rc == SpB_SUCCESS)
        for (int64_t t = 0; t < P; t++) nec[t] = (int32_t)(win_off[t + 1] - win_off[t]);
    vnec_free_dev(dco);
    vnec_free_dev(decr);
    vnec_free_dev(dwin);
    vnec_free_dev(doff);
    return rc;
}

/* VNEC-S task order and reordered matrix on the device at (P, L, cut), over
 * the snapped partitions (spmv_vnec_S_for_mxv.c:336-585), copied to the host
 * buffers given (any of s_val, task_start, task_end may be NULL). */
SpB_Info vnec_s_order_ext(const SpB_Matrix A, int64_t P, int L, int cut, int32_t *tasks,
                          int32_t *spvv_len, int32_t *row_begin, int32_t *row_end, double *s_val,
                          int32_t *s_col, int32_t *task_start, int32_t *task_end)
{
    spb_csr *a = spb_as_csr(A);
    if (!a || P < 1 || L < 1 || !tasks || !spvv_len || !row_begin || !row_end || !s_col)
        return vnec_fail("vnec_s_order_ext: bad arguments");
    if (spb_matrix_upload(a) != SpB_SUCCESS) return SpB_FAILURE;
    hipStream_t s = vnec_stream();
    int64_t nr = (int64_t)a->base.row, nnz = (int64_t)a->nnz;
    size_t nrb = sizeof(int32_t) * (size_t)(nr ? nr : 1), nzb = (size_t)(nnz ? nnz : 1);
    int32_t *dco = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * 2 * (size_t)(P + 1), NULL);
    int32_t *decr = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * nzb, NULL);
    int32_t *dwin = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * nzb, NULL);
    int64_t *doff = (int64_t *)vnec_malloc_dev(sizeof(int64_t) * (size_t)(P + 1), NULL);
    int32_t *dtasks = (int32_t *)vnec_malloc_dev(nrb, NULL);
    int32_t *dspvv = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)P, NULL);
    int32_t *drb = (int32_t *)vnec_malloc_dev(nrb, NULL);
    int32_t *dre = (int32_t *)vnec_malloc_dev(nrb, NULL);
    int32_t *dcol = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * nzb, NULL);
    double *dval = s_val ? (double *)vnec_malloc_dev(sizeof(double) * nzb, NULL) : NULL;
    int32_t *hco = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(P + 1));
    SpB_Info rc = (dco && decr && dwin && doff && dtasks && dspvv && drb && dre && dcol && hco &&
                   (!s_val || dval)) ? SpB_SUCCESS : vnec_fail("out of memory");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_partition(a->d_ptr, nr, nnz, P, 1, dco, s), "partition");
    if (rc == SpB_SUCCESS) rc = ecr_device(a, P, dco, 0, decr, dwin, doff);
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_s_order(a->d_ptr, nr, dco, P, L, cut, dtasks, dspvv, s), "s order");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_s_reorder(a->d_ptr, a->d_val, decr, dtasks, nr, nnz, dco, dspvv, P, L, drb, dre,
                                            dval, dcol, s), "s reorder");
    if (rc == SpB_SUCCESS && nr) {
        rc = vnec_hip_ok(hipMemcpyAsync(tasks, dtasks, nrb, hipMemcpyDeviceToHost, s), "d2h");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemcpyAsync(row_begin, drb, nrb, hipMemcpyDeviceToHost, s), "d2h");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemcpyAsync(row_end, dre, nrb, hipMemcpyDeviceToHost, s), "d2h");
    }
    if (rc == SpB_SUCCESS && nnz) {
        rc = vnec_hip_ok(hipMemcpyAsync(s_col, dcol, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost, s), "d2h");
        if (rc == SpB_SUCCESS && s_val)
            rc = vnec_hip_ok(hipMemcpyAsync(s_val, dval, sizeof(double) * (size_t)nnz, hipMemcpyDeviceToHost, s), "d2h");
    }
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(spvv_len, dspvv, sizeof(int32_t) * (size_t)P, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(hco, dco, sizeof(int32_t) * 2 * (size_t)(P + 1), hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    /* tasks of partition t occupy [start_t.x, end_t.x) (snapped rows are contiguous) */
    for (int64_t t = 0; rc == SpB_SUCCESS && t < P; t++) {
        if (task_start) task_start[t] = hco[2 * t];
        if (task_end) task_end[t] = hco[2 * t + 2];
    }
    vnec_free_dev(dco);
    vnec_free_dev(decr);
    vnec_free_dev(dwin);
    vnec_free_dev(doff);
    vnec_free_dev(dtasks);
    vnec_free_dev(dspvv);
    vnec_free_dev(drb);
    vnec_free_dev(dre);
    vnec_free_dev(dcol);
    vnec_free_dev(dval);
    free(hco);
    return rc;
}

/* D layout on the device at (P, L).  Device outputs: coords, ecr, win,
 * win_off, v_row_ptr (nrows + 1), col_start (G), val_align (G * L, only
 * when want_vals).  *G receives the group count. */
typedef struct {
    int32_t *coords, *ecr, *win, *vrp, *col_start;
    int64_t *win_off;
    double *val_align;
    int64_t G;
} dlay;

static void dlay_free(dlay *d)
{
    vnec_free_dev(d->coords);
    vnec_free_dev(d->ecr);
    vnec_free_dev(d->win);
    vnec_free_dev(d->vrp);
    vnec_free_dev(d->col_start);
    vnec_free_dev(d->win_off);
    vnec_free_dev(d->val_align);
    memset(d, 0, sizeof *d);
}

static SpB_Info d_layout_device(spb_csr *a, int64_t P, int L, int want_vals, dlay *d)
{
    memset(d, 0, sizeof *d);
    if (spb_matrix_upload(a) != SpB_SUCCESS) return SpB_FAILURE;
    hipStream_t s = vnec_stream();
    int64_t nr = (int64_t)a->base.row, nnz = (int64_t)a->nnz;
    d->coords = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * 2 * (size_t)(P + 1), NULL);
    d->ecr = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nnz ? nnz : 1), NULL);
    d->win = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nnz ? nnz : 1), NULL);
    d->win_off = (int64_t *)vnec_malloc_dev(sizeof(int64_t) * (size_t)(P + 1), NULL);
    d->vrp = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nr + 1), NULL);
    SpB_Info rc = (d->coords && d->ecr && d->win && d->win_off && d->vrp) ? SpB_SUCCESS : SpB_FAILURE;
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_partition(a->d_ptr, nr, nnz, P, 0, d->coords, s), "partition");
    if (rc == SpB_SUCCESS) rc = ecr_device(a, P, d->coords, 0, d->ecr, d->win, d->win_off);
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_d_count(a->d_ptr, d->ecr, nr, d->coords, P, L, d->vrp, NULL, s), "d count");
    int64_t G = 0;
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_scan_excl_i32(d->vrp, d->vrp, nr + 1, &G, s), "d scan");
    if (rc == SpB_SUCCESS) {
        d->G = G;
        d->col_start = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(G ? G : 1), NULL);
        if (!d->col_start) rc = SpB_FAILURE;
        if (want_vals && rc == SpB_SUCCESS) {
            d->val_align = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(G > 0 ? G * L : 1), NULL);
            if (!d->val_align) rc = SpB_FAILURE;
            else rc = vnec_hip_ok(hipMemsetAsync(d->val_align, 0, sizeof(double) * (size_t)(G * L), s), "zero");
        }
    }
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_d_fill(a->d_ptr, d->ecr, a->d_val, nr, d->coords, P, L, d->vrp,
                                         d->col_start, d->val_align, 0, s), "d fill");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    if (rc != SpB_SUCCESS) dlay_free(d);
    return rc;
}

/* Group count only (what IRD needs). */
SpB_Info vnec_d_layout_ext(const SpB_Matrix A, int64_t P, int L, int32_t *v_row_ptr,
                           int32_t *col_start, double *val_align, int64_t *groups)
{
    spb_csr *a = spb_as_csr(A);
    if (!a || P < 1 || L < 1) return vnec_fail("vnec_d_layout_ext: bad arguments");
    dlay d;
    if (d_layout_device(a, P, L, val_align != NULL, &d) != SpB_SUCCESS) return SpB_FAILURE;
    hipStream_t s = vnec_stream();
    int64_t nr = (int64_t)a->base.row;
    SpB_Info rc = SpB_SUCCESS;
    if (groups) *groups = d.G;
    if (v_row_ptr)
        rc = vnec_hip_ok(hipMemcpyAsync(v_row_ptr, d.vrp, sizeof(int32_t) * (size_t)(nr + 1), hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS && col_start && d.G)
        rc = vnec_hip_ok(hipMemcpyAsync(col_start, d.col_start, sizeof(int32_t) * (size_t)d.G, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS && val_align && d.G)
        rc = vnec_hip_ok(hipMemcpyAsync(val_align, d.val_align, sizeof(double) * (size_t)(d.G * L), hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    dlay_free(&d);
    return rc;
}

/* --------------------------------------------------------------- IRD */
/* ird_calc.c:2-123 — IRD = (float)nnz / (float)(G * L) at P = 36, L = 8,
 * D when IRD >= (float)0.477. */
SpB_Info vnec_ird_ext(const SpB_Matrix A, int64_t P, int L, double thr, float *ird, int *decision)
{
    spb_csr *a = spb_as_csr(A);
    if (!a) return SpB_FAILURE;
    int64_t G = 0;
    if (vnec_d_layout_ext(A, P, L, NULL, NULL, NULL, &G) != SpB_SUCCESS) return SpB_FAILURE;
    int32_t slots = (int32_t)(G * L);
    float r = (float)a->nnz / (float)slots;
    if (ird) *ird = r;
    if (decision) *decision = r >= (float)thr ? SpB_VNEC_D : SpB_VNEC_S;
    return SpB_SUCCESS;
}

SpB_VNEC_type IRD_VNEC(const SpB_Matrix A, SpB_Type type)
{
    int dec = SpB_VNEC_S;
    vnec_clear_error(); /* no status return: the error slot tells */
    if (type != SpB_FP64 && type != SpB_FP32) {
        vnec_fail("IRD_VNEC: type must be SpB_FP64 or SpB_FP32");
        return SpB_VNEC_S;
    }
    int L = type == SpB_FP32 ? REF_L32 : REF_L;
    double thr = type == SpB_FP32 ? REF_IRD_THR32 : REF_IRD_THR;
    if (vnec_ird_ext(A, REF_P, L, thr, NULL, &dec) != SpB_SUCCESS) return SpB_VNEC_S;
    return (SpB_VNEC_type)dec;
}

/* ------------------------------------------------------ public plans */
/* The device plan behind a Build_VNEC_* result runs the matching GPU
 * kernel: VNEC-L a merge-path plan (for FP64 the learned selector's pick
 * when it is one, else 7 items per thread, persistent for FP64, per-tile
 * carries for FP32; tiles in column order where the column-reuse rule
 * holds), VNEC-D the wave64
 * D-block kernel (groups of the reference's nLanes window positions: 8 for
 * FP64, 16 for FP32), VNEC-S the 64-row slice kernel when the slice layout
 * applies (padding <= 3) and otherwise the learned selector's choice
 * (FP64) or the merge-path kernel (FP32). */
static struct vnec_plan_s *gpu_plan_for(spb_csr *a, plan_kind kind)
{
    vnec_plan p = NULL;
    const int fp32 = a->base.element_type == SpB_FP32;
    vnec_tuning t = {VNEC_SELECT_LEARNED, 7, 64, fp32 ? VNEC_ENC_FP32 : VNEC_ENC_FP64, 1, -1};
    if (fp32) t.variant = VNEC_VAR_L; /* grid -1: one workgroup per tile */
    if (kind == KIND_L) {
        /* merge-path CSR: for FP64 the learned selector's pick when it is a
         * merge-path plan (CSR-ordered or column-sorted tiles, and beyond the
         * model's training sizes the fastest of the forms measured around
         * it); otherwise tiles stored in column order where columns are
         * reused (the same algorithm, fewer lines per gather) */
        vnec_tuning sel;
        if (!fp32 && vnec_select_tuning(a, VNEC_SELECT_LEARNED, 1, &sel) == SpB_SUCCESS &&
            (sel.variant == VNEC_VAR_L || sel.variant == VNEC_VAR_C)) {
            vnec_tuning lt = {VNEC_SELECT_LEARNED, 7, 64, VNEC_ENC_FP64, 1, -1};
            if (vnec_plan_build_ext(&p, (SpB_Matrix)a, &lt) != SpB_SUCCESS) return NULL;
            return p;
        }
        t.variant = vnec_csort_rule(a) ? VNEC_VAR_C : VNEC_VAR_L;
        t.grid = fp32 ? -1 : 0;
    } else if (vnec_pattern_rule(a)) {
        /* rows of one column-offset pattern: pattern slices, the slice
         * layout without column indices.  This serves the D entry too: IRD
         * sends matrices of dense column runs (banded, block-diagonal) to
         * D, whose blocks the GPU streams at about half the rate of the
         * slice layouts on them */
        t.variant = VNEC_VAR_G;
        t.ipt = 8;
    } else {
        uint32_t mask = 0;
        int slices = 0;
        if (vnec_candidate_mask_ext((SpB_Matrix)a, &mask) != SpB_SUCCESS) return NULL;
        vnec_tuning c[VNEC_NCAND];
        vnec_candidates_ext(c, VNEC_NCAND);
        for (int i = 0; i < VNEC_NCAND; i++)
            if (c[i].variant == VNEC_VAR_S && ((mask >> i) & 1u)) slices = 1;
        if (slices) {
            t.variant = VNEC_VAR_S;
            t.ipt = 3;
        } else if (kind == KIND_D32) {
            /* FP32 D blocks (the FP32 API has no selector) */
            t.variant = VNEC_VAR_D;
            t.lanes = REF_L32;
        }
        /* FP64 otherwise: the learned selector, whose candidates include
         * the D blocks at widths 4, 8 and 16 */
    }
    if (vnec_plan_build_ext(&p, (SpB_Matrix)a, &t) != SpB_SUCCESS) return NULL;
    return p;
}

/* spmv_vnec_L_for_mxv.c:2-35 (one plan type for both value types) */
struct VNEC_L *Build_VNEC_L(const SpB_Matrix A)
{
    spb_csr *a = spb_as_csr(A);
    if (!a) return NULL;
    int64_t nr = (int64_t)a->base.row, nnz = (int64_t)a->nnz;
    VNEC_L *l = (VNEC_L *)calloc(1, sizeof(VNEC_L));
    int32_t *co = (int32_t *)malloc(sizeof(int32_t) * 2 * (REF_P + 1));
    if (!l || !co) {
        free(l);
        free(co);
        vnec_fail("Build_VNEC_L: out of memory");
        return NULL;
    }
    if (vnec_partition_ext(A, REF_P, 0, co) != SpB_SUCCESS) {
        free(l);
        free(co);
        return NULL;
    }
    int64_t ipt = ipt_of(nr, nnz, REF_P), M = nr + nnz;
    l->num_merge_items = (int)M;
    l->items_per_thread = (int)ipt;
    l->diagonal_start = (int *)malloc(sizeof(int) * REF_P);
    l->diagonal_end = (int *)malloc(sizeof(int) * REF_P);
    l->thread_coord_start = (coord *)malloc(sizeof(coord) * REF_P);
    l->thread_coord_end = (coord *)malloc(sizeof(coord) * REF_P);
    l->nz_indices = (int *)malloc(sizeof(int) * (size_t)(nnz ? nnz : 1));
    if (!l->diagonal_start || !l->diagonal_end || !l->thread_coord_start || !l->thread_coord_end ||
        !l->nz_indices) {
        free(co);
        VNEC_L_free_ext(l);
        vnec_fail("Build_VNEC_L: out of memory");
        return NULL;
    }
    for (int t = 0; t < REF_P; t++) {
        int64_t ds = ipt * t < M ? ipt * t : M;
        l->diagonal_start[t] = (int)ds;
        l->diagonal_end[t] = (int)(ds + ipt < M ? ds + ipt : M);
        l->thread_coord_start[t].x = co[2 * t];
        l->thread_coord_start[t].y = co[2 * t + 1];
        l->thread_coord_end[t].x = co[2 * t + 2];
        l->thread_coord_end[t].y = co[2 * t + 3];
    }
    for (int64_t j = 0; j < nnz; j++) l->nz_indices[j] = (int)j;
    free(co);
    struct vnec_plan_s *g = gpu_plan_for(a, KIND_L);
    if (!g) {
        VNEC_L_free_ext(l);
        return NULL;
    }
    add_wrap(l, KIND_L, g, a);
    return l;
}

/* Host fields the D plans of both value types share (all but the compact
 * x buffers ecr_xx_val, whose element type differs). */
typedef struct {
    coord *tcs, *tce;
    int *ecr_indices, *NEC_NUM, **use_x_indices, *v_row_ptr, *col_start;
    double *val_align; /* G * L, FP64 (exact for FP32 matrices) */
    int64_t G;
} d_fields;

static void d_fields_free(d_fields *f)
{
    for (int t = 0; f->use_x_indices && t < REF_P; t++) free(f->use_x_indices[t]);
    free(f->tcs);
    free(f->tce);
    free(f->ecr_indices);
    free(f->NEC_NUM);
    free(f->use_x_indices);
    free(f->v_row_ptr);
    free(f->col_start);
    free(f->val_align);
    memset(f, 0, sizeof *f);
}

/* 64-byte aligned host array of n bytes (SpB_Alloc.c:53-66 aligns the
 * reference's arrays the same way); released with free(). */
static void *host_aligned(size_t n)
{
    void *p = NULL;
    if (posix_memalign(&p, 64, n ? n : 64) != 0) return NULL;
    return p;
}

/* spmv_vnec_D_for_mxv.c:242-398 (FP64) and :2-241 (FP32): partitions,
 * ECR, the group layout and val_align (spmv_vnec_D_for_mxv.c:345-392:
 * group g's values at [g * L, (g + 1) * L), zero-padded, a repeated
 * column's last value kept) at P = 36 and L lanes, computed on the GPU. */
static SpB_Info d_public(spb_csr *a, int L, const char *fn, d_fields *f)
{
    memset(f, 0, sizeof *f);
    int64_t nr = (int64_t)a->base.row, nnz = (int64_t)a->nnz;
    dlay d;
    if (d_layout_device(a, REF_P, L, 1, &d) != SpB_SUCCESS) return SpB_FAILURE;
    f->G = d.G;
    int32_t *co = (int32_t *)malloc(sizeof(int32_t) * 2 * (REF_P + 1));
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (REF_P + 1));
    int32_t *win = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    hipStream_t s = vnec_stream();
    f->tcs = (coord *)malloc(sizeof(coord) * REF_P);
    f->tce = (coord *)malloc(sizeof(coord) * REF_P);
    f->ecr_indices = (int *)calloc((size_t)nnz + 10, sizeof(int));
    f->NEC_NUM = (int *)calloc(REF_P + 1, sizeof(int));
    f->use_x_indices = (int **)calloc(REF_P + 1, sizeof(int *));
    f->v_row_ptr = (int *)malloc(sizeof(int) * (size_t)(nr + 1));
    f->col_start = (int *)malloc(sizeof(int) * (size_t)(d.G ? d.G : 1));
    f->val_align = (double *)host_aligned(sizeof(double) * (size_t)(d.G * L));
    SpB_Info rc = (co && off && win && f->tcs && f->tce && f->ecr_indices && f->NEC_NUM &&
                   f->use_x_indices && f->v_row_ptr && f->col_start && f->val_align)
                      ? SpB_SUCCESS
                      : vnec_fail("%s: out of memory", fn);
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(co, d.coords, sizeof(int32_t) * 2 * (REF_P + 1), hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(off, d.win_off, sizeof(int64_t) * (REF_P + 1), hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS && nnz)
        rc = vnec_hip_ok(hipMemcpyAsync(f->ecr_indices, d.ecr, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(f->v_row_ptr, d.vrp, sizeof(int32_t) * (size_t)(nr + 1), hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS && d.G)
        rc = vnec_hip_ok(hipMemcpyAsync(f->col_start, d.col_start, sizeof(int32_t) * (size_t)d.G, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS && d.G)
        rc = vnec_hip_ok(hipMemcpyAsync(f->val_align, d.val_align, sizeof(double) * (size_t)(d.G * L),
                                        hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS && nnz)
        rc = vnec_hip_ok(hipMemcpyAsync(win, d.win, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    for (int t = 0; rc == SpB_SUCCESS && t < REF_P; t++) {
        f->tcs[t].x = co[2 * t];
        f->tcs[t].y = co[2 * t + 1];
        f->tce[t].x = co[2 * t + 2];
        f->tce[t].y = co[2 * t + 3];
        int n = (int)(off[t + 1] - off[t]);
        f->NEC_NUM[t] = n;
        f->use_x_indices[t] = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
        if (!f->use_x_indices[t]) rc = vnec_fail("%s: out of memory", fn);
        else memcpy(f->use_x_indices[t], win + off[t], sizeof(int) * (size_t)n);
    }
    free(co);
    free(off);
    free(win);
    dlay_free(&d);
    if (rc != SpB_SUCCESS) d_fields_free(f);
    return rc;
}

struct VNEC_D_FP64 *Build_VNEC_D_FP64(const SpB_Matrix A)
{
    spb_csr *a = spb_as_csr(A);
    if (!a) return NULL;
    if (a->base.element_type != SpB_FP64) {
        vnec_fail("Build_VNEC_D_FP64: the matrix holds FP32 values (use Build_VNEC_D_FP32)");
        return NULL;
    }
    d_fields f;
    if (d_public(a, REF_L, "Build_VNEC_D_FP64", &f) != SpB_SUCCESS) return NULL;
    VNEC_D_FP64 *p = (VNEC_D_FP64 *)calloc(1, sizeof(VNEC_D_FP64));
    double **xx = (double **)calloc(REF_P + 1, sizeof(double *));
    SpB_Info rc = (p && xx) ? SpB_SUCCESS : vnec_fail("Build_VNEC_D_FP64: out of memory");
    for (int t = 0; rc == SpB_SUCCESS && t < REF_P; t++) {
        xx[t] = (double *)malloc(sizeof(double) * (size_t)(f.NEC_NUM[t] + REF_L));
        if (!xx[t]) rc = vnec_fail("Build_VNEC_D_FP64: out of memory");
    }
    if (rc != SpB_SUCCESS) {
        for (int t = 0; xx && t < REF_P; t++) free(xx[t]);
        free(xx);
        free(p);
        d_fields_free(&f);
        return NULL;
    }
    p->thread_coord_start = f.tcs;
    p->thread_coord_end = f.tce;
    p->ecr_indices = f.ecr_indices;
    p->NEC_NUM = f.NEC_NUM;
    p->use_x_indices = f.use_x_indices;
    p->ecr_xx_val = xx;
    p->v_row_ptr = f.v_row_ptr;
    p->col_start = f.col_start;
    p->val_align = f.val_align;
    struct vnec_plan_s *g = gpu_plan_for(a, KIND_D);
    if (!g) {
        VNEC_D_free_fp64_ext(p);
        return NULL;
    }
    add_wrap(p, KIND_D, g, a);
    return p;
}

struct VNEC_D_FP32 *Build_VNEC_D_FP32(const SpB_Matrix A)
{
    spb_csr *a = spb_as_csr(A);
    if (!a) return NULL;
    if (a->base.element_type != SpB_FP32) {
        vnec_fail("Build_VNEC_D_FP32: the matrix holds FP64 values (use Build_VNEC_D_FP64)");
        return NULL;
    }
    d_fields f;
    if (d_public(a, REF_L32, "Build_VNEC_D_FP32", &f) != SpB_SUCCESS) return NULL;
    VNEC_D_FP32 *p = (VNEC_D_FP32 *)calloc(1, sizeof(VNEC_D_FP32));
    float **xx = (float **)calloc(REF_P + 1, sizeof(float *));
    SpB_Info rc = (p && xx) ? SpB_SUCCESS : vnec_fail("Build_VNEC_D_FP32: out of memory");
    for (int t = 0; rc == SpB_SUCCESS && t < REF_P; t++) {
        xx[t] = (float *)malloc(sizeof(float) * (size_t)(f.NEC_NUM[t] + REF_L32));
        if (!xx[t]) rc = vnec_fail("Build_VNEC_D_FP32: out of memory");
    }
    if (rc != SpB_SUCCESS) {
        for (int t = 0; xx && t < REF_P; t++) free(xx[t]);
        free(xx);
        free(p);
        d_fields_free(&f);
        return NULL;
    }
    p->thread_coord_start = f.tcs;
    p->thread_coord_end = f.tce;
    p->ecr_indices = f.ecr_indices;
    p->NEC_NUM = f.NEC_NUM;
    p->use_x_indices = f.use_x_indices;
    p->ecr_xx_val = xx;
    p->v_row_ptr = f.v_row_ptr;
    p->col_start = f.col_start;
    /* the device layout holds the FP32 values widened, so narrowing is exact */
    size_t nva = (size_t)(f.G * REF_L32);
    p->val_align = (float *)host_aligned(sizeof(float) * nva);
    if (!p->val_align) {
        vnec_fail("Build_VNEC_D_FP32: out of memory");
        free(f.val_align);
        VNEC_D_free_fp32_ext(p);
        return NULL;
    }
    for (size_t i = 0; i < nva; i++) p->val_align[i] = (float)f.val_align[i];
    free(f.val_align);
    struct vnec_plan_s *g = gpu_plan_for(a, KIND_D32);
    if (!g) {
        VNEC_D_free_fp32_ext(p);
        return NULL;
    }
    add_wrap(p, KIND_D32, g, a);
    return p;
}

/* Host fields the S plans of both value types share; val holds the
 * reordered values in FP64 (exact for FP32 matrices, whose device mirror
 * is their exact widening) for the typed builder to keep or narrow. */
typedef struct {
    int *NEC_NUM, **use_x_indices, *spvv_len, **tasks;
    int *col, *row_begin, *row_end, *task_start, *task_end;
    double *val;
} s_fields;

static void s_fields_free(s_fields *f)
{
    for (int t = 0; t < REF_P; t++) {
        if (f->use_x_indices) free(f->use_x_indices[t]);
        if (f->tasks) free(f->tasks[t]);
    }
    free(f->NEC_NUM);
    free(f->use_x_indices);
    free(f->spvv_len);
    free(f->tasks);
    free(f->col);
    free(f->row_begin);
    free(f->row_end);
    free(f->task_start);
    free(f->task_end);
    free(f->val);
    memset(f, 0, sizeof *f);
}

/* spmv_vnec_S_for_mxv.c:336-588 (FP64) and :3-264 (FP32): snapped
 * partitions, their ECR, the task order (rows of length <= 32 in groups of
 * L first) and the reordered matrix, at P = 36 and L lanes. */
static SpB_Info s_public(spb_csr *a, int L, const char *fn, s_fields *f)
{
    memset(f, 0, sizeof *f);
    int64_t nnz = (int64_t)a->nnz, nr = (int64_t)a->base.row;
    int32_t *co = (int32_t *)malloc(sizeof(int32_t) * 2 * (REF_P + 1));
    int32_t *ecr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    int32_t *nec = (int32_t *)malloc(sizeof(int32_t) * REF_P);
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (REF_P + 1));
    int32_t *win = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    int32_t *tasks = NULL;
    size_t rb = sizeof(int) * (size_t)(nr ? nr : 1), zb = (size_t)(nnz ? nnz : 1);
    SpB_Info rc = (co && ecr && nec && off && win) ? SpB_SUCCESS : vnec_fail("%s: out of memory", fn);
    if (rc == SpB_SUCCESS) rc = vnec_partition_ext((SpB_Matrix)a, REF_P, 1, co);
    if (rc == SpB_SUCCESS) rc = vnec_ecr_ext((SpB_Matrix)a, REF_P, co, 0, ecr, nec, off, win);
    if (rc == SpB_SUCCESS) {
        f->NEC_NUM = (int *)calloc(REF_P + 1, sizeof(int));
        f->use_x_indices = (int **)calloc(REF_P + 1, sizeof(int *));
        f->spvv_len = (int *)calloc(REF_P + 1, sizeof(int));
        f->tasks = (int **)calloc(REF_P + 1, sizeof(int *));
        f->val = (double *)malloc(sizeof(double) * zb);
        f->col = (int *)malloc(sizeof(int) * zb);
        f->row_begin = (int *)malloc(rb);
        f->row_end = (int *)malloc(rb);
        f->task_start = (int *)calloc(REF_P + 1, sizeof(int));
        f->task_end = (int *)calloc(REF_P + 1, sizeof(int));
        tasks = (int32_t *)malloc(rb);
        if (!f->NEC_NUM || !f->use_x_indices || !f->spvv_len || !f->tasks || !f->val || !f->col ||
            !f->row_begin || !f->row_end || !f->task_start || !f->task_end || !tasks)
            rc = vnec_fail("%s: out of memory", fn);
    }
    for (int t = 0; rc == SpB_SUCCESS && t < REF_P; t++) {
        int n = nec[t];
        f->NEC_NUM[t] = n;
        f->use_x_indices[t] = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
        if (!f->use_x_indices[t]) rc = vnec_fail("%s: out of memory", fn);
        else memcpy(f->use_x_indices[t], win + off[t], sizeof(int) * (size_t)n);
    }
    /* task order and reordered matrix (spmv_vnec_S_for_mxv.c:442-585) */
    if (rc == SpB_SUCCESS)
        rc = vnec_s_order_ext((SpB_Matrix)a, REF_P, L, REF_CROSS_MAX, tasks, f->spvv_len, f->row_begin,
                              f->row_end, f->val, f->col, f->task_start, f->task_end);
    for (int t = 0; rc == SpB_SUCCESS && t < REF_P; t++) {
        int n = f->task_end[t] - f->task_start[t];
        f->tasks[t] = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        if (!f->tasks[t]) rc = vnec_fail("%s: out of memory", fn);
        else if (n > 0) memcpy(f->tasks[t], tasks + f->task_start[t], sizeof(int) * (size_t)n);
    }
    free(tasks);
    free(co);
    free(ecr);
    free(nec);
    free(off);
    free(win);
    if (rc != SpB_SUCCESS) s_fields_free(f);
    return rc;
}

static double **alloc_xx_f64(const int *nec)
{
    double **xx = (double **)calloc(REF_P + 1, sizeof(double *));
    for (int t = 0; xx && t < REF_P; t++)
        if (!(xx[t] = (double *)malloc(sizeof(double) * (size_t)(nec[t] ? nec[t] : 1)))) {
            for (int u = 0; u < t; u++) free(xx[u]);
            free(xx);
            return NULL;
        }
    return xx;
}

static float **alloc_xx_f32(const int *nec)
{
    float **xx = (float **)calloc(REF_P + 1, sizeof(float *));
    for (int t = 0; xx && t < REF_P; t++)
        if (!(xx[t] = (float *)malloc(sizeof(float) * (size_t)(nec[t] ? nec[t] : 1)))) {
            for (int u = 0; u < t; u++) free(xx[u]);
            free(xx);
            return NULL;
        }
    return xx;
}

struct VNEC_S_FP64 *Build_VNEC_S_FP64(const SpB_Matrix A)
{
    spb_csr *a = spb_as_csr(A);
    if (!a) return NULL;
    if (a->base.element_type != SpB_FP64) {
        vnec_fail("Build_VNEC_S_FP64: the matrix holds FP32 values (use Build_VNEC_S_FP32)");
        return NULL;
    }
    s_fields f;
    if (s_public(a, REF_L, "Build_VNEC_S_FP64", &f) != SpB_SUCCESS) return NULL;
    VNEC_S_FP64 *p = (VNEC_S_FP64 *)calloc(1, sizeof(VNEC_S_FP64));
    double **xx = alloc_xx_f64(f.NEC_NUM);
    if (!p || !xx) {
        free(p);
        if (xx)
            for (int t = 0; t < REF_P; t++) free(xx[t]);
        free(xx);
        s_fields_free(&f);
        vnec_fail("Build_VNEC_S_FP64: out of memory");
        return NULL;
    }
    p->NEC_NUM = f.NEC_NUM;
    p->use_x_indices = f.use_x_indices;
    p->ecr_xx_val = xx;
    p->spvv8_len = f.spvv_len;
    p->tasks = f.tasks;
    p->reorder_mat.nnz = f.val;
    p->reorder_mat.col = f.col;
    p->reorder_mat.row_begin = f.row_begin;
    p->reorder_mat.row_end = f.row_end;
    p->reorder_mat.task_start = f.task_start;
    p->reorder_mat.task_end = f.task_end;
    struct vnec_plan_s *g = gpu_plan_for(a, KIND_S);
    if (!g) {
        VNEC_S_free_fp64_ext(p);
        return NULL;
    }
    add_wrap(p, KIND_S, g, a);
    return p;
}

struct VNEC_S_FP32 *Build_VNEC_S_FP32(const SpB_Matrix A)
{
    spb_csr *a = spb_as_csr(A);
    if (!a) return NULL;
    if (a->base.element_type != SpB_FP32) {
        vnec_fail("Build_VNEC_S_FP32: the matrix holds FP64 values (use Build_VNEC_S_FP64)");
        return NULL;
    }
    s_fields f;
    if (s_public(a, REF_L32, "Build_VNEC_S_FP32", &f) != SpB_SUCCESS) return NULL;
    int64_t nnz = (int64_t)a->nnz;
    VNEC_S_FP32 *p = (VNEC_S_FP32 *)calloc(1, sizeof(VNEC_S_FP32));
    float **xx = alloc_xx_f32(f.NEC_NUM);
    float *v32 = (float *)malloc(sizeof(float) * (size_t)(nnz ? nnz : 1));
    if (!p || !xx || !v32) {
        free(p);
        if (xx)
            for (int t = 0; t < REF_P; t++) free(xx[t]);
        free(xx);
        free(v32);
        s_fields_free(&f);
        vnec_fail("Build_VNEC_S_FP32: out of memory");
        return NULL;
    }
    for (int64_t j = 0; j < nnz; j++) v32[j] = (float)f.val[j]; /* exact: widened FP32 values */
    free(f.val);
    f.val = NULL;
    p->NEC_NUM = f.NEC_NUM;
    p->use_x_indices = f.use_x_indices;
    p->ecr_xx_val = xx;
    p->spvv16_len = f.spvv_len;
    p->tasks = f.tasks;
    p->reorder_mat.nnz = v32;
    p->reorder_mat.col = f.col;
    p->reorder_mat.row_begin = f.row_begin;
    p->reorder_mat.row_end = f.row_end;
    p->reorder_mat.task_start = f.task_start;
    p->reorder_mat.task_end = f.task_end;
    struct vnec_plan_s *g = gpu_plan_for(a, KIND_S32);
    if (!g) {
        VNEC_S_free_fp32_ext(p);
        return NULL;
    }
    add_wrap(p, KIND_S32, g, a);
    return p;
}

/* ------------------------------------------------------------- SpMV */
static SpB_Info check_types(const spb_csr *a, SpB_Vector y, SpB_Vector x, SpB_Type want, const char *fn)
{
    spb_dense *xv = spb_as_dense(x), *yv = spb_as_dense(y);
    if (!xv || !yv) return SpB_FAILURE;
    const char *tn = want == SpB_FP32 ? "FP32" : "FP64";
    if (a->base.element_type != want) return vnec_fail("%s: the matrix does not hold %s values", fn, tn);
    if (xv->base.element_type != want || yv->base.element_type != want)
        return vnec_fail("%s: x and y must be %s vectors", fn, tn);
    return SpB_SUCCESS;
}

static SpB_Info run_wrapped(const void *pub, plan_kind kind, SpB_Type type, SpB_Vector y,
                            const SpB_Matrix A, const SpB_Vector x, const char *fn)
{
    vnec_clear_error();
    wrap *w = find_wrap(pub);
    if (!w || w->kind != kind) return vnec_fail("%s: plan was not made by this library's Build_VNEC_*", fn);
    if ((const spb_csr *)A != w->A) return vnec_fail("%s: plan was built for another matrix", fn);
    if (check_types(w->A, y, (SpB_Vector)x, type, fn) != SpB_SUCCESS) return SpB_FAILURE;
    return vnec_spmv_ext(w->gpu, y, x);
}

void SpMV_VNEC_L_FP64(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_L *mat_thd)
{
    run_wrapped(mat_thd, KIND_L, SpB_FP64, y, A, x, "SpMV_VNEC_L_FP64");
}

void SpMV_VNEC_L_FP32(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_L *mat_thd)
{
    run_wrapped(mat_thd, KIND_L, SpB_FP32, y, A, x, "SpMV_VNEC_L_FP32");
}

void SpMV_VNEC_D_FP64(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_D_FP64 *mat_thd)
{
    run_wrapped(mat_thd, KIND_D, SpB_FP64, y, A, x, "SpMV_VNEC_D_FP64");
}

void SpMV_VNEC_D_FP32(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_D_FP32 *mat_thd)
{
    run_wrapped(mat_thd, KIND_D32, SpB_FP32, y, A, x, "SpMV_VNEC_D_FP32");
}

void SpMV_VNEC_S_FP64(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_S_FP64 *mat_thd)
{
    run_wrapped(mat_thd, KIND_S, SpB_FP64, y, A, x, "SpMV_VNEC_S_FP64");
}

void SpMV_VNEC_S_FP32(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_S_FP32 *mat_thd)
{
    run_wrapped(mat_thd, KIND_S32, SpB_FP32, y, A, x, "SpMV_VNEC_S_FP32");
}

SpB_Info SpB_mxv(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, void *VNEC)
{
    spb_csr *a = spb_as_csr(A);
    if (!a) return SpB_FAILURE;
    spb_dense *xv = spb_as_dense(x), *yv = spb_as_dense(y);
    if (!xv || !yv) return SpB_FAILURE;
    if (xv->base.n != a->base.col || yv->base.n != a->base.row)
        return vnec_fail("SpB_mxv: dimension mismatch (A %lux%lu, x %lu, y %lu)", a->base.row,
                         a->base.col, xv->base.n, yv->base.n);
    if (check_types(a, y, x, a->base.element_type, "SpB_mxv") != SpB_SUCCESS) return SpB_FAILURE;
    if (VNEC) {
        wrap *w = find_wrap(VNEC);
        if (w) {
            if (w->A != a) return vnec_fail("SpB_mxv: plan was built for another matrix");
            return vnec_spmv_ext(w->gpu, y, x);
        }
        struct vnec_plan_s *p = (struct vnec_plan_s *)VNEC;
        if (p->magic != SPB_MAGIC_PLAN) return vnec_fail("SpB_mxv: unknown plan handle");
        if (p->A != a) return vnec_fail("SpB_mxv: plan was built for another matrix");
        return vnec_spmv_ext(p, y, x);
    }
    if (!a->auto_plan && vnec_plan_build_ext(&a->auto_plan, A, NULL) != SpB_SUCCESS) return SpB_FAILURE;
    return vnec_spmv_ext(a->auto_plan, y, x);
}

/* ---------------------------------------------------------- freeing */
SpB_Info VNEC_L_free_ext(VNEC_L *l)
{
    if (!l) return vnec_fail("VNEC_L_free_ext: NULL plan");
    struct vnec_plan_s *g = drop_wrap(l);
    if (g) vnec_plan_free_ext(&g);
    free(l->diagonal_start);
    free(l->diagonal_end);
    free(l->thread_coord_start);
    free(l->thread_coord_end);
    free(l->nz_indices);
    free(l);
    return SpB_SUCCESS;
}

SpB_Info VNEC_D_free_fp64_ext(VNEC_D_FP64 *p)
{
    if (!p) return vnec_fail("VNEC_D_free_fp64_ext: NULL plan");
    struct vnec_plan_s *g = drop_wrap(p);
    if (g) vnec_plan_free_ext(&g);
    for (int t = 0; p->ecr_xx_val && t < REF_P; t++) free(p->ecr_xx_val[t]);
    free(p->ecr_xx_val);
    d_fields f = {p->thread_coord_start, p->thread_coord_end, p->ecr_indices, p->NEC_NUM,
                  p->use_x_indices, p->v_row_ptr, p->col_start};
    d_fields_free(&f);
    free(p->val_align);
    free(p);
    return SpB_SUCCESS;
}

SpB_Info VNEC_D_free_fp32_ext(VNEC_D_FP32 *p)
{
    if (!p) return vnec_fail("VNEC_D_free_fp32_ext: NULL plan");
    struct vnec_plan_s *g = drop_wrap(p);
    if (g) vnec_plan_free_ext(&g);
    for (int t = 0; p->ecr_xx_val && t < REF_P; t++) free(p->ecr_xx_val[t]);
    free(p->ecr_xx_val);
    d_fields f = {p->thread_coord_start, p->thread_coord_end, p->ecr_indices, p->NEC_NUM,
                  p->use_x_indices, p->v_row_ptr, p->col_start};
    d_fields_free(&f);
    free(p->val_align);
    free(p);
    return SpB_SUCCESS;
}

SpB_Info VNEC_S_free_fp64_ext(VNEC_S_FP64 *p)
{
    if (!p) return vnec_fail("VNEC_S_free_fp64_ext: NULL plan");
    struct vnec_plan_s *g = drop_wrap(p);
    if (g) vnec_plan_free_ext(&g);
    for (int t = 0; p->ecr_xx_val && t < REF_P; t++) free(p->ecr_xx_val[t]);
    free(p->ecr_xx_val);
    s_fields f = {p->NEC_NUM, p->use_x_indices, p->spvv8_len, p->tasks, p->reorder_mat.col,
                  p->reorder_mat.row_begin, p->reorder_mat.row_end, p->reorder_mat.task_start,
                  p->reorder_mat.task_end, p->reorder_mat.nnz};
    s_fields_free(&f);
    free(p);
    return SpB_SUCCESS;
}

SpB_Info VNEC_S_free_fp32_ext(VNEC_S_FP32 *p)
{
    if (!p) return vnec_fail("VNEC_S_free_fp32_ext: NULL plan");
    struct vnec_plan_s *g = drop_wrap(p);
    if (g) vnec_plan_free_ext(&g);
    for (int t = 0; p->ecr_xx_val && t < REF_P; t++) free(p->ecr_xx_val[t]);
    free(p->ecr_xx_val);
    free(p->reorder_mat.nnz);
    s_fields f = {p->NEC_NUM, p->use_x_indices, p->spvv16_len, p->tasks, p->reorder_mat.col,
                  p->reorder_mat.row_begin, p->reorder_mat.row_end, p->reorder_mat.task_start,
                  p->reorder_mat.task_end, NULL};
    s_fields_free(&f);
    free(p);
    return SpB_SUCCESS;
}
