/*
 * vnec_ext.h — extensions of libvnec_mi355x.so beyond the reference API.
 *
 * None of these has a counterpart in the reference header unless a line
 * says so; they exist for callers that need what a GPU library adds:
 * device-resident vectors, explicit plan lifetimes, GPU execution plans
 * with tunable tiles, value encodings, multi-right-hand-side and batched
 * launches, the builder structures at arbitrary parameters (for parity
 * checks), timing on the library's stream, and the variant selector.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "OpenSparseBLAS.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reason for the last SpB_FAILURE / NULL return on this thread. */
const char *vnec_last_error(void);
/* Number of visible GPUs (0 when the HIP runtime finds none). */
int vnec_device_count_ext(void);

/* ------------------------------------------------------------ objects */
/* Fill a matrix made by SpB_Matrix_new (CSR) straight from CSR arrays
 * (ptr: nrows + 1 entries, columns sorted within each row).  Equivalent to
 * SpB_Matrix_build_FP64 on the same entries, without the sort. */
SpB_Info SpB_Matrix_import_csr_ext(SpB_Matrix A, const int64_t *ptr, const int32_t *col,
                                   const double *val);
SpB_Info SpB_Matrix_nvals_ext(SpB_Index *nvals, const SpB_Matrix A);
/* Batch of nmat independent matrices as one block-diagonal CSR matrix
 * (A made by SpB_Matrix_new with row_off[nmat] rows, col_off[nmat]
 * columns): member m owns rows [row_off[m], row_off[m+1]) and columns
 * [col_off[m], col_off[m+1]); col holds member-local columns.  A plan on A
 * runs the whole batch, each member with its own x slice, in one launch. */
SpB_Info vnec_batch_import_ext(SpB_Matrix A, int64_t nmat, const int64_t *row_off,
                               const int64_t *col_off, const int64_t *ptr, const int32_t *col,
                               const double *val);
/* Bulk copy of a dense FP64 vector's n values in / out. */
SpB_Info SpB_Vector_import_ext(SpB_Vector v, const double *values);
SpB_Info SpB_Vector_export_ext(const SpB_Vector v, double *values);
/* Device residency: when on, SpMV reads x from / writes y to device memory
 * only; host values are refreshed by SpB_Vector_sync_host_ext or lazily by
 * extractElement / export, and setElement / import mark the device copy
 * stale.  Off (the default) keeps the reference's host semantics. */
SpB_Info SpB_Vector_device_resident_ext(SpB_Vector v, bool on);
SpB_Info SpB_Vector_sync_host_ext(SpB_Vector v);
/* Device address of a vector's values (allocated on first use). */
double *SpB_Vector_device_ptr_ext(SpB_Vector v);

/* --------------------------------------------------------- GPU plans */
typedef enum {
    VNEC_VAR_L = 0, /* merge-path CSR, direct x gather, int32 columns       */
    VNEC_VAR_E = 1, /* merge-path CSR, ECR x window staged in LDS, u16 cols */
    VNEC_VAR_S = 2, /* row-aligned tiles, 64-row equal-length slices        */
    VNEC_VAR_D = 3, /* dense column blocks of width `lanes`                 */
    VNEC_VAR_N
} vnec_variant;

typedef enum {
    VNEC_ENC_FP64 = 0, /* FP64 values                                   */
    VNEC_ENC_HILO = 1, /* FP32 hi + FP32 lo, a = hi + lo, FP64 accumulate */
    VNEC_ENC_HI = 2,   /* FP32 hi only, FP64 accumulate                 */
    VNEC_ENC_N
} vnec_encoding;

typedef struct {
    int variant;  /* vnec_variant, or -1: choose with the learned selector */
    int ipt;      /* merge items per thread (tile = 256 * ipt items)       */
    int lanes;    /* D block width / S slice height                        */
    int encoding; /* vnec_encoding                                         */
    int nrhs;     /* right-hand sides per SpMV (1, or 2..16)               */
    int grid;     /* 0: persistent, one workgroup per resident slot;
                     > 0: persistent with this many workgroups (multiple
                     of 8); -1: one workgroup per tile                     */
} vnec_tuning;

typedef struct vnec_plan_s *vnec_plan;

/* Build a GPU execution plan for A.  t may be NULL (selector picks). */
SpB_Info vnec_plan_build_ext(vnec_plan *plan, const SpB_Matrix A, const vnec_tuning *t);
SpB_Info vnec_plan_free_ext(vnec_plan *plan);
/* Tuning in force, tile count and device bytes held by the plan. */
SpB_Info vnec_plan_info_ext(vnec_plan plan, vnec_tuning *t, int64_t *ntiles, int64_t *dev_bytes);

/* y <- A x (x: ncols * nrhs, y: nrows * nrhs, row-major [i][rhs]).
 * Synchronous: y is valid when the call returns. */
SpB_Info vnec_spmv_ext(vnec_plan plan, SpB_Vector y, const SpB_Vector x);
/* Same on caller device buffers, asynchronous on the library stream. */
SpB_Info vnec_spmv_dev_ext(vnec_plan plan, const double *x_dev, double *y_dev);
/* Times `iters` back-to-back SpMVs on the library stream with HIP events
 * (optionally replayed from one captured hipGraph); *ms = mean ms per SpMV.
 * If flush_bytes > 0 a buffer of that size is overwritten before each SpMV
 * (outside the timed kernels: *ms then excludes it, using per-SpMV event
 * pairs) so that the matrix cannot stay resident in the on-die caches. */
SpB_Info vnec_time_dev_ext(vnec_plan plan, const double *x_dev, double *y_dev, int iters,
                           int use_graph, size_t flush_bytes, double *ms);

/* Mean duration (ms) of the SpMV kernel and of the carry fix-up kernel,
 * each timed between its own HIP events on the library stream. */
SpB_Info vnec_time_kernels_ext(vnec_plan plan, const double *x_dev, double *y_dev, int iters,
                               double *ms_main, double *ms_fixup);

/* Measured streaming-copy bandwidth: bytes read + bytes written per
 * second over iters copies of a `bytes`-sized buffer (GB/s). */
SpB_Info vnec_copy_bandwidth_ext(size_t bytes, int iters, double *gbps);

/* ------------------------------------------------- device memory helpers */
void *vnec_dev_alloc_ext(size_t bytes);
SpB_Info vnec_dev_free_ext(void *p);
SpB_Info vnec_memcpy_h2d_ext(void *dst, const void *src, size_t bytes);
SpB_Info vnec_memcpy_d2h_ext(void *dst, const void *src, size_t bytes);
SpB_Info vnec_sync_ext(void);

/* ----------------------------- builder structures at any parameters
 * These run the GPU builders and copy their output to host buffers, so the
 * structures can be checked bit-for-bit against the CPU partitioner and
 * builders at the reference parameters (P = 36, L = 8) or the GPU ones. */

/* Merge-path coordinates of the P + 1 diagonals min(t * ceil(M / P), M),
 * t = 0..P (coords: 2 (P + 1) ints, (x, y) pairs).  snap != 0 applies the
 * S-variant row snap.  Replaces MergePathDivide + the diagonal loop
 * (spmv_vnec_common.c:3-33, spmv_vnec_D_for_mxv.c:250-268). */
SpB_Info vnec_partition_ext(const SpB_Matrix A, int64_t P, int snap, int32_t *coords);

/* ECR of the partitions given by coords (2 (P + 1) ints): ecr (nnz),
 * nec (P), win_off (P + 1), win (sum nec <= nnz).  method 0 = automatic,
 * 1 = dense flag scan, 2 = per-tile LDS sort.  Replaces
 * spmv_vnec_D_for_mxv.c:280-338. */
SpB_Info vnec_ecr_ext(const SpB_Matrix A, int64_t P, const int32_t *coords, int method,
                      int32_t *ecr, int32_t *nec, int64_t *win_off, int32_t *win);

/* VNEC-D block layout at (P, L): v_row_ptr (nrows + 1), col_start (G),
 * val_align (G * L, may be NULL), *groups = G.  Any output may be NULL.
 * Replaces the grouping loop of spmv_vnec_D_for_mxv.c:356-392. */
SpB_Info vnec_d_layout_ext(const SpB_Matrix A, int64_t P, int L, int32_t *v_row_ptr,
                           int32_t *col_start, double *val_align, int64_t *groups);

/* IRD ratio (float)nnz / (float)(G * L) and the D / S decision at (P, L,
 * threshold); IRD_VNEC uses P = 36, L = 8, 0.477 (ird_calc.c:2-123). */
SpB_Info vnec_ird_ext(const SpB_Matrix A, int64_t P, int L, double thr, float *ird, int *decision);

/* ------------------------------------------------------- plan lifetimes
 * The reference has no destroy call for its plans; these release both the
 * public host fields and the device plan behind them. */
SpB_Info VNEC_D_free_fp64_ext(VNEC_D_FP64 *plan);
SpB_Info VNEC_S_free_fp64_ext(VNEC_S_FP64 *plan);
SpB_Info VNEC_L_free_ext(VNEC_L *plan);
/* Device plan behind a Build_VNEC_* result (owned by it). */
vnec_plan VNEC_plan_of_ext(const void *reference_plan);

#ifdef __cplusplus
}
#endif
