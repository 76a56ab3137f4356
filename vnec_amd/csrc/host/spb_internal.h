/*
 * spb_internal.h — private object layouts of libvnec_mi355x.so.
 *
 * The matrix and vector structs start with the reference's own layouts
 * (src/core/include/SpB_Matrix.h:10-17 + SpB_Matrix_CSC_or_CSR.h:12-21,
 * SpB_Vector.h:10-15 + SpB_Vector_Dense.h:15-19), so code that downcasts a
 * handle the way the reference's operation layer and test harness do keeps
 * reading the same fields.  Private fields (device mirrors, residency
 * flags, cached plans) follow the reference prefix.
 */
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdint.h>

#include "OpenSparseBLAS.h"
#include "vnec_ext.h"
#include "vnec_hip.h"

#define SPB_MAGIC_MATRIX 0x4D58564Eu /* "NVXM" */
#define SPB_MAGIC_VECTOR 0x5658564Eu
#define SPB_MAGIC_PLAN 0x5050564Eu

struct SpB_Matrix_opaque {
    SpB_Type element_type;
    SpB_Matrix_Format_Value_ext format_type;
    SpB_Index row;
    SpB_Index col;
    size_t element_size;
};

struct SpB_Vector_opaque {
    SpB_Type element_type;
    SpB_Vector_Format_Value_ext format_type;
    SpB_Index n;
    size_t element_size;
};

typedef struct spb_csr {
    struct SpB_Matrix_opaque base;
    void *val;          /* nnz values (host), FP64 or FP32   */
    SpB_Index *indices; /* nnz column indices (host, 64-bit) */
    SpB_Index *ptr;     /* ptr_len row offsets (host)        */
    SpB_Index ptr_len;
    SpB_Index nnz;
    SpB_Index capacity;
    /* ---- private ---- */
    uint32_t magic;
    int dev_ready;      /* device mirror matches the host arrays */
    int32_t *d_ptr;     /* nrows + 1 */
    int32_t *d_col;     /* nnz       */
    double *d_val;      /* nnz, FP64 (FP32 values widened)   */
    struct vnec_plan_s *auto_plan; /* plan cached by SpB_mxv(..., NULL) */
    int tuned_cand;     /* autotuner's best candidate, -1 until measured */
    int tuned_fp32;     /* FP32 autotuner's best shape, -1 until measured */
    int refined_ipt;    /* merge-path neighbourhood measured at this ipt
                           (0: not yet): the fastest of L / C under
                           both schedules                            */
    int refined_variant, refined_grid, refined_lanes, refined_ipt_out;
    int refined_slice_ipt; /* slice-family pick measured at this ipt (0: not
                              yet): the fastest of S at it and variant G  */
    int refined_slice_variant, refined_slice_ipt_out;
    int feat_ready;     /* feat_ints holds the selector's statistics     */
    int pattern_ready;  /* pattern_rows holds the count below            */
    int64_t pattern_rows; /* rows that fit the dominant column-offset pattern */
    uint64_t feat_ints[VNEC_NFEAT_INTS];
} spb_csr;

typedef struct spb_dense {
    struct SpB_Vector_opaque base;
    void *values; /* n values (host), FP64 or FP32 */
    /* ---- private ---- */
    uint32_t magic;
    void *d_values; /* n values (device), same type */
    int resident;   /* device copy is authoritative between calls      */
    int host_stale; /* device holds newer values than host             */
    int dev_stale;  /* host holds newer values than device             */
    int pinned;     /* values come from hipHostMalloc (page-locked)    */
} spb_dense;

/* Captured SpMV graphs a plan keeps (distinct x / y device pointer pairs). */
#define VNEC_PLAN_GRAPHS 4

struct vnec_plan_s {
    uint32_t magic;
    vnec_tuning tune;
    int64_t nrows, ncols, nnz;
    int64_t ntiles;
    int32_t *coords;   /* 2 (ntiles + 1) */
    double *carry;     /* ntiles * nrhs  (one workgroup per tile)      */
    int nwg;           /* persistent workgroups (0: one per tile)       */
    int snwg;          /* > 0: per-tile carries from the strided
                          persistent launch of this many workgroups  */
    int32_t *wg_row;   /* nwg: row of each workgroup's trailing carry   */
    double *wg_carry;  /* nwg                                           */
    /* variant E */
    uint16_t *lcol;
    int32_t *run_off, *run_start, *run_wpos, *nec;
    /* encodings */
    float *hi, *lo;
    /* variant S: 64-row slices */
    int64_t nslices;
    int32_t *slice_len, *slice_ptr, *tile_slices;
    double *sv;
    float *shi, *slo;
    int32_t *scol;
    uint16_t *slcol;
    int32_t max_nec;
    int win;
    int64_t s_padded;   /* variant S: slots of the padded slice layout     */
    int64_t nruns;      /* window runs over all tiles (variants E, D, S)   */
    /* variant D: groups of `lanes` window positions */
    int32_t *vrp;       /* nrows + 1 */
    int32_t *col_start; /* groups    */
    double *dvals;      /* groups * lanes, wave64 layout */
    float *dvals32;     /* the same in FP32 (encoding FP32) */
    int64_t groups;
    /* variant H: hot x entries */
    int32_t *hcol;      /* nnz: column, or -1 - slot for a hot column */
    int32_t *hot_order; /* ncols: columns by descending use; first nhot are hot */
    double *xhot;       /* nhot: packed before each SpMV */
    int nhot;
    /* variant G: pattern slices + a sub-plan for the irregular rows */
    int32_t *g_pat;     /* L offsets                                  */
    int32_t g_pat_h[VNEC_PATTERN_MAX]; /* the same offsets on the host */
    int g_L;
    uint32_t *g_masks;  /* nrows                                      */
    int64_t g_nslices;
    double *g_vals;     /* nslices * L * 64 (FP64 encoding)           */
    float *g_hi, *g_lo; /* the same slots, FP32 hi (+ lo) encodings   */
    int32_t *g_irr;     /* nirr rows that are not pattern subsets     */
    int64_t g_nirr;
    SpB_Matrix g_sub;   /* CSR of the irregular rows (owned)          */
    struct vnec_plan_s *g_subplan;
    double *g_ysub;     /* nirr (floats for FP32 plans)               */
    /* variant C: each tile's nonzeros in column order */
    double *cs_val;     /* nnz                                        */
    int32_t *cs_col;    /* nnz                                        */
    uint16_t *cs_pos;   /* nnz: offset within the tile's CSR range,
                           NULL when packed into cs_col's high bits  */
    int cs_shift;       /* > 0: cs_col = column | offset << cs_shift  */
    int32_t *cs_base;   /* ntiles, or NULL: with cs_shift > 0, the packed
                           column is relative to its tile's smallest one  */
    /* relabelled columns (variants L and C, lanes = VNEC_LANES_RELABEL):
       label h is column cperm[h] (columns by descending use); rcol holds the
       relabelled columns, xperm the permuted x of the current SpMV */
    int32_t *cperm;     /* ncols */
    int32_t *rcol;      /* nnz   */
    double *xperm;      /* ncols */
    /* borrowed from the matrix mirror */
    const int32_t *ptr, *col;
    const double *val;
    const spb_csr *A;
    int64_t dev_bytes;
    /* cached graph for timing */
    hipGraphExec_t gexec;
    /* captured single SpMVs for replay (vnec_spmv_graph_ext), one per
       (x, y) device pointer pair, least recently used replaced */
    struct {
        hipGraphExec_t exec;
        const void *x;
        void *y;
        uint64_t used;
    } sgraph[VNEC_PLAN_GRAPHS];
    uint64_t sgraph_clock;
    /* interleaved copies of column-major right-hand sides (multi-RHS) */
    double *xt, *yt;
    const void *g_x;
    void *g_y;
    int g_reps;
};

/* ---- library runtime (spb_core.c) ---- */
hipStream_t vnec_stream(void);
SpB_Info vnec_fail(const char *fmt, ...);
/* Entry points without a status return (the reference's SpMV_VNEC_*,
 * IRD_VNEC) clear the error slot first, so vnec_last_error() afterwards
 * describes that call. */
void vnec_clear_error(void);
SpB_Info vnec_hip_ok(int rc, const char *what);
int vnec_runtime_ready(void);
void *vnec_malloc_dev(size_t bytes, int64_t *acct);
void vnec_free_dev(void *p);
spb_csr *spb_as_csr(const SpB_Matrix A);
spb_dense *spb_as_dense(const SpB_Vector v);
SpB_Info spb_matrix_upload(spb_csr *a);
SpB_Info spb_vector_to_device(spb_dense *v);
SpB_Info spb_vector_from_device(spb_dense *v);
SpB_Info spb_vector_device_alloc(spb_dense *v);

/* ---- variant G (spb_pattern.c) ---- */
/* Dominant row pattern of a (host arrays): its offsets and length (0: no
 * row qualifies), picked among evenly spaced sample rows. */
int vnec_pattern_choose(const spb_csr *a, int32_t pat[VNEC_PATTERN_MAX]);
/* Builds the pattern slices and the irregular-row sub-plan of p. */
SpB_Info vnec_pattern_build(struct vnec_plan_s *p, spb_csr *a);
/* Rows of a whose columns are a subset of its dominant pattern (0 when
 * there is none). */
SpB_Info vnec_pattern_coverage(spb_csr *a, int64_t *nregular);

/* ---- plans (spb_plan.c) ---- */
/* x_dev / y_dev: doubles, or floats for an encoding-FP32 plan. */
SpB_Info vnec_plan_run(struct vnec_plan_s *p, const void *x_dev, void *y_dev);
double vnec_slice_padding(const spb_csr *a);

/* ---- autotuner and selector (spb_tune.c) ---- */
/* Tuning chosen by a selection mode (vnec_select_mode) for FP64 values
 * and nrhs right-hand sides (several: the merge-path kernel, its
 * granularity autotuned in VNEC_SELECT_AUTOTUNE mode, else 5 items per
 * thread). */
SpB_Info vnec_select_tuning(spb_csr *a, int mode, int nrhs, vnec_tuning *t);
/* The learned selector's pick among every candidate but pattern slices
 * (the plan for the rows a pattern plan leaves over). */
SpB_Info vnec_select_tuning_no_pattern(spb_csr *a, vnec_tuning *t);
/* Times the FP64 merge-path plan t (variant L or C) in its four forms —
 * CSR-ordered (L) or column-sorted (C) tiles, persistent (grid 0) or
 * per-tile carries (grid -1) — at t's granularity, and, when x takes at
 * least 2 MiB, the same four with columns relabelled by descending use;
 * column-sorted 512-thread tiles at t's granularity (9 above it),
 * persistent, with per-tile carries, and (x >= 2 MiB) with per-tile
 * carries over relabelled columns; and 512-thread tiles with their gathers
 * issued together at t's granularity (7 above it); sets t's variant, ipt,
 * grid and lanes to the fastest.  Cached on the matrix. */
SpB_Info vnec_refine_merge_plan(spb_csr *a, vnec_tuning *t);
/* The autotune mode on an FP32 matrix: every FP32 plan shape that applies
 * (64-row slices, pattern slices, L and C with per-tile carries at several
 * granularities, C also in 512-thread tiles, with and without gathers issued
 * together, D at block width 16) built and timed on the matrix; t is set to
 * the fastest.  Cached on the matrix. */
SpB_Info vnec_autotune_fp32(spb_csr *a, vnec_tuning *t);
/* The structural rules of the learned and rule modes: most rows fit one
 * column-offset pattern (variant G applies), and tiles reuse columns
 * (variant C beats L).  FP64 matrices. */
int vnec_pattern_rule(spb_csr *a);
/* Times the FP64 slice-family plan t (variant S, or G) as 64-row slices at
 * t's granularity (3 items per thread for G) and as pattern slices where
 * they apply, and sets t to the faster.  Cached on the matrix. */
SpB_Info vnec_refine_slice_plan(spb_csr *a, vnec_tuning *t);
/* Fills g's x-window fields for pattern pat (L ascending offsets): the
 * merged column runs a workgroup's 256 rows read, staged in LDS when they
 * total at most VNEC_PATTERN_WIN_MAX entries (else wsize = 0). */
void vnec_pattern_window(const int32_t *pat, int L, vnec_pattern_args *g);
int vnec_csort_rule(spb_csr *a);

/* ---- compiled-in selector model (spb_selector_model.c, generated) ---- */
extern const int vnec_model_trained;
extern const int64_t vnec_model_train_samples;
extern const float vnec_model_params[VNEC_MLP_NPARAMS];
extern const float vnec_model_mu[VNEC_NFEAT];
extern const float vnec_model_sd[VNEC_NFEAT];
