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
/* The same for a matrix made with SpB_FP32 (SpB_Matrix_build_FP32 order). */
SpB_Info SpB_Matrix_import_csr_fp32_ext(SpB_Matrix A, const int64_t *ptr, const int32_t *col,
                                        const float *val);
SpB_Info SpB_Matrix_nvals_ext(SpB_Index *nvals, const SpB_Matrix A);
/* Batch of nmat independent matrices as one block-diagonal CSR matrix
 * (A made by SpB_Matrix_new with row_off[nmat] rows, col_off[nmat]
 * columns): member m owns rows [row_off[m], row_off[m+1]) and columns
 * [col_off[m], col_off[m+1]); col holds member-local columns.  A plan on A
 * runs the whole batch, each member with its own x slice, in one launch. */
SpB_Info vnec_batch_import_ext(SpB_Matrix A, int64_t nmat, const int64_t *row_off,
                               const int64_t *col_off, const int64_t *ptr, const int32_t *col,
                               const double *val);
/* Graph files as the reference harness reads them (utilities/reader.hpp,
 * builder.hpp): .mtx (1-based coordinates; symmetric files add reversed
 * edges; sorted, duplicate edges and self-loops dropped) or .sg (serialized
 * CSR, taken as stored).  ptr (nrows + 1) and col are allocated here; free
 * them with vnec_host_free_ext. */
SpB_Info vnec_read_graph_ext(const char *path, int64_t *nrows, int64_t **ptr, int32_t **col);
void vnec_host_free_ext(void *p);
/* A square CSR matrix (type SpB_FP64 or SpB_FP32) from such a file, with
 * the values the reference harness gives the pattern
 * (test_spmv_vnec.cpp:160-170: a_ij = i, its known-answer test) or all
 * ones.  Selects SpB_CSR as the matrix format. */
enum { VNEC_VALUES_ROW = 0, VNEC_VALUES_ONE = 1 };
SpB_Info SpB_Matrix_read_ext(SpB_Matrix *A, const char *path, SpB_Type type, int values);

/* Bulk copy of a dense FP64 vector's n values in / out. */
SpB_Info SpB_Vector_import_ext(SpB_Vector v, const double *values);
SpB_Info SpB_Vector_export_ext(const SpB_Vector v, double *values);
/* The same for an FP32 vector. */
SpB_Info SpB_Vector_import_fp32_ext(SpB_Vector v, const float *values);
SpB_Info SpB_Vector_export_fp32_ext(const SpB_Vector v, float *values);
/* Device residency: when on, SpMV reads x from / writes y to device memory
 * only; host values are refreshed by SpB_Vector_sync_host_ext or lazily by
 * extractElement / export, and setElement / import mark the device copy
 * stale.  Off (the default) keeps the reference's host semantics. */
SpB_Info SpB_Vector_device_resident_ext(SpB_Vector v, bool on);
SpB_Info SpB_Vector_sync_host_ext(SpB_Vector v);
/* Device address of a vector's values (allocated on first use; FP64 or
 * FP32 elements, as the vector was made). */
void *SpB_Vector_device_ptr_ext(SpB_Vector v);

/* --------------------------------------------------------- GPU plans */
typedef enum {
    VNEC_VAR_L = 0, /* merge-path CSR, direct x gather, int32 columns       */
    VNEC_VAR_E = 1, /* merge-path CSR, ECR x window staged in LDS, u16 cols */
    VNEC_VAR_S = 2, /* row-aligned tiles, 64-row equal-length slices        */
    VNEC_VAR_D = 3, /* dense column blocks of width `lanes`                 */
    VNEC_VAR_H = 4, /* persistent merge-path CSR with the `lanes` (4096 or
                       8192) most used x entries held in LDS                */
    VNEC_VAR_G = 5, /* pattern slices: rows whose columns are a subset of
                       the dominant column-offset pattern store a position
                       mask and values only; the rest run a sub-plan       */
    VNEC_VAR_C = 6, /* persistent merge-path CSR whose tiles store their
                       nonzeros in column order (+ its offset in the
                       tile): x is gathered in ascending column order;
                       FP64, one right-hand side                          */
    VNEC_VAR_N
} vnec_variant;

typedef enum {
    VNEC_ENC_FP64 = 0, /* FP64 values                                   */
    VNEC_ENC_HILO = 1, /* FP32 hi + FP32 lo, a = hi + lo, FP64 accumulate */
    VNEC_ENC_HI = 2,   /* FP32 hi only, FP64 accumulate                 */
    VNEC_ENC_FP32 = 3, /* the FP32 API: FP32 values, x, y and sums (FP32
                          matrices only; variants L and C (per-tile
                          carries: grid -1 or -2), S, D and G, one
                          right-hand side)                              */
    VNEC_ENC_N
} vnec_encoding;

/* Negative vnec_tuning.variant values: how the plan chooses its variant,
 * tile granularity and schedule. */
typedef enum {
    VNEC_SELECT_LEARNED = -1,   /* the trained MLP over matrix statistics (the
                                   rule below when no model is installed)    */
    VNEC_SELECT_AUTOTUNE = -2,  /* exhaustive: build and time every candidate
                                   (FP32 matrices: every FP32 plan shape)    */
    VNEC_SELECT_RULE = -3,      /* S when 64-row slice padding <= 1.15, else
                                   persistent L at 7 items per thread        */
    VNEC_SELECT_REFERENCE = -4, /* the reference dispatch (nrows >= 3e6: L;
                                   else IRD at P = 36, L = 8: D -> E, S -> S) */
} vnec_select_mode;

typedef struct {
    int variant;  /* vnec_variant, or a vnec_select_mode                    */
    int ipt;      /* merge items per thread (tile = 256 * ipt items)       */
    int lanes;    /* D block width / H hot x entries in LDS; L and C:
                     64, or the flags VNEC_LANES_RELABEL (columns
                     relabelled by descending use, x permuted before each
                     SpMV) and, for C, VNEC_LANES_WIDE (512-thread
                     tiles) and VNEC_LANES_GATHER (gathers issued
                     together), alone or or-ed together                  */
    int encoding; /* vnec_encoding                                         */
    int nrhs;     /* right-hand sides per SpMV: 1, 2, 4, 8 or 16           */
    int grid;     /* 0: persistent, one workgroup per resident slot;
                     > 0: persistent with this many workgroups (multiple
                     of 8); -1: every tile leaves its own carry, one
                     workgroup per tile (or, with at least as many
                     tiles as resident workgroups, and always with
                     VNEC_LANES_WIDE, the form of -2); -2: the same
                     per-tile carries from a persistent launch whose
                     workgroups sweep each XCD's tiles together (variants
                     L and C, one right-hand side)                       */
} vnec_tuning;

/* vnec_tuning.lanes of an L or C plan whose columns are relabelled by
 * descending use: the most used x entries share cache lines (the same sums
 * in the same order as without it). */
#define VNEC_LANES_RELABEL 1
/* vnec_tuning.lanes of a C plan whose tiles hold 512 * ipt items (ipt 3,
 * 5, 7, 9 or 11; one right-hand side; 512-thread workgroups, FP64 from a
 * persistent launch or with per-tile carries, FP32 with per-tile carries)
 * instead of 256 * ipt: each
 * tile's nonzeros are sorted by column over the larger window, so the x
 * gathers of a wavefront share more cache lines. */
#define VNEC_LANES_WIDE 1024
/* vnec_tuning.lanes flag of a C plan with per-tile carries (FP64 or FP32,
 * ipt 3, 5 or 7, one right-hand side; alone or with VNEC_LANES_WIDE): each thread
 * issues all of a tile's x gathers before using any, instead of waiting for
 * each in turn.  More loads in flight per wavefront, for more registers:
 * faster where x sits in the caches (batches of small matrices), slower
 * where the gathers reach HBM. */
#define VNEC_LANES_GATHER 2048

typedef struct vnec_plan_s *vnec_plan;

/* Build a GPU execution plan for A.  t may be NULL (selector picks). */
SpB_Info vnec_plan_build_ext(vnec_plan *plan, const SpB_Matrix A, const vnec_tuning *t);
SpB_Info vnec_plan_free_ext(vnec_plan *plan);
/* Tuning in force, tile count and device bytes held by the plan. */
SpB_Info vnec_plan_info_ext(vnec_plan plan, vnec_tuning *t, int64_t *ntiles, int64_t *dev_bytes);
/* Bytes the plan's own layout must move per SpMV: every array its kernels
 * read counted once, x read once and y written once (padding slots
 * included, caches ignored).  *main_bytes: the main SpMV kernel alone;
 * *spmv_bytes: the whole SpMV (variant G adds its irregular-row sub-plan).
 * Either may be NULL.  The format-agnostic figure, CSR with int32 columns,
 * is (vb + 4) nnz + 4 (nrows + 1) + 8 k (ncols + nrows). */
SpB_Info vnec_plan_bytes_ext(vnec_plan plan, int64_t *main_bytes, int64_t *spmv_bytes);

/* y <- A x (x: ncols * nrhs, y: nrows * nrhs, row-major [i][rhs]).
 * Synchronous: y is valid when the call returns. */
SpB_Info vnec_spmv_ext(vnec_plan plan, SpB_Vector y, const SpB_Vector x);
/* Same on caller device buffers, asynchronous on the library stream.
 * Device buffers hold doubles, or floats for VNEC_ENC_FP32 plans.  With
 * nrhs >= 4, x and y 16-byte aligned run the faster grouped multi-RHS
 * kernel; 8-byte aligned ones are accepted and run the one-lane kernel. */
SpB_Info vnec_spmv_dev_ext(vnec_plan plan, const void *x_dev, void *y_dev);
/* Same, replayed from a HIP graph: the plan captures its launch sequence
 * for this (x_dev, y_dev) pair on the first call and replays it after
 * (bit-identical to vnec_spmv_dev_ext; x_dev's contents may change between
 * calls, its address may not).  A plan keeps graphs for 4 pairs; a fifth
 * recaptures over the least recently used.  One replay costs more host and
 * launch time than the plan's one to three direct kernel launches, so
 * vnec_spmv_dev_ext, vnec_spmv_ext and SpB_mxv launch directly. */
SpB_Info vnec_spmv_graph_ext(vnec_plan plan, const void *x_dev, void *y_dev);
/* Same with the right-hand sides column-major: X holds nrhs vectors of
 * ncols doubles one after another, Y nrhs vectors of nrows.  The plan
 * interleaves X into scratch of its own (allocated on the first call),
 * runs, and writes Y back column-major; for nrhs = 1 this is
 * vnec_spmv_dev_ext. */
SpB_Info vnec_spmm_colmajor_dev_ext(vnec_plan plan, const void *X_dev, void *Y_dev);
/* Times `iters` back-to-back SpMVs on the library stream with HIP events
 * (optionally replayed from one captured hipGraph); *ms = mean ms per SpMV.
 * If flush_bytes > 0 a buffer of that size (make it larger than the L2 and
 * Infinity Cache together) is read with cache-allocating loads before each
 * SpMV, so that the matrix cannot stay resident in the on-die caches; *ms
 * is then the time of iters (flush, SpMV) pairs less that of iters flushes
 * alone, each sequence back to back between one pair of events. */
SpB_Info vnec_time_dev_ext(vnec_plan plan, const void *x_dev, void *y_dev, int iters,
                           int use_graph, size_t flush_bytes, double *ms);

/* Mean duration (ms) of the SpMV kernel and of the carry fix-up kernel:
 * iters launches of each, back to back between one pair of HIP events on the
 * library stream.  y holds A·x on return. */
SpB_Info vnec_time_kernels_ext(vnec_plan plan, const void *x_dev, void *y_dev, int iters,
                               double *ms_main, double *ms_fixup);

/* Measured streaming-copy bandwidth: bytes read + bytes written per
 * second over iters copies of a `bytes`-sized buffer (GB/s). */
SpB_Info vnec_copy_bandwidth_ext(size_t bytes, int iters, double *gbps);
/* Measured streaming-read bandwidth: bytes read per second over iters
 * reads of a `bytes`-sized buffer (GB/s); an SpMV is read-dominated. */
SpB_Info vnec_read_bandwidth_ext(size_t bytes, int iters, double *gbps);
/* Measured ceilings of A's SpMV traffic without row work (a balanced
 * persistent grid over its nonzeros): mean ms to stream its FP64 values
 * and int32 columns, and to stream them and gather x_dev[col] for every
 * nonzero (iters launches each, HIP events). */
SpB_Info vnec_gather_ceiling_ext(const SpB_Matrix A, const double *x_dev, int iters, double *stream_ms,
                                 double *gather_ms);

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

/* VNEC-S task order and reordered matrix at (P, L, cross-row cut) over the
 * snapped partitions (reference: P = 36, L = 8, cut 32): tasks (nrows,
 * concatenated per partition), spvv_len (P), row_begin / row_end (nrows),
 * s_val (nnz, may be NULL), s_col (nnz, ECR ranks), task_start / task_end
 * (P).  Replaces spmv_vnec_S_for_mxv.c:442-585. */
SpB_Info vnec_s_order_ext(const SpB_Matrix A, int64_t P, int L, int cut, int32_t *tasks,
                          int32_t *spvv_len, int32_t *row_begin, int32_t *row_end, double *s_val,
                          int32_t *s_col, int32_t *task_start, int32_t *task_end);

/* Variant G's row pattern: the column offsets P[0 .. L) (position l of
 * row i is column i + P[l]) of the most frequent row among 32 sample rows
 * i_s = ((s + 1) 2654435761 mod 2^32) nrows / 2^32 (the first on ties);
 * *L = 0 when no sample row has 1 .. VNEC_PATTERN_MAX entries.
 * *nregular receives the rows whose columns are a subset of the pattern
 * (counted on the GPU); the plan stores those as masks and values. */
#define VNEC_PATTERN_MAX 32
SpB_Info vnec_pattern_ext(const SpB_Matrix A, int32_t pattern[VNEC_PATTERN_MAX], int *L,
                          int64_t *nregular);
/* masks[i] = positions of the pattern row i of A holds, flags[i] = 1 when
 * it is not a subset of the pattern (GPU; host arrays of nrows). */
SpB_Info vnec_pattern_match_ext(const SpB_Matrix A, const int32_t *pattern, int L, uint32_t *masks,
                                int32_t *flags);

/* IRD ratio (float)nnz / (float)(G * L) and the D / S decision at (P, L,
 * threshold); IRD_VNEC uses P = 36 and L = 8, 0.477 for SpB_FP64 or
 * L = 16, 0.437 for SpB_FP32 (ird_calc.c:2-123, SpB_mxv.h:26-33). */
SpB_Info vnec_ird_ext(const SpB_Matrix A, int64_t P, int L, double thr, float *ird, int *decision);

/* --------------------------------------------- autotuner and selector
 * No counterpart in the reference (its only selector is IRD_VNEC plus the
 * row-count gate of its test harness); these serve callers that tune,
 * train or audit the variant choice. */

/* Candidate configurations (single right-hand side, FP64).  Copies up to
 * max of them to out (may be NULL); returns VNEC_NCAND. */
#define VNEC_NCAND 23
int vnec_candidates_ext(vnec_tuning *out, int max);
/* Which candidates may run on A: bit c set when candidate c applies
 * (slice candidates need padding <= 3 and fewer than 2^31 slots; D
 * candidates need 8 * lanes * nnz <= 2^32 bytes of value blocks; the
 * pattern-slice candidate (21) needs at least half of the rows on the
 * dominant column-offset pattern and fewer than 2^31 slots). */
SpB_Info vnec_candidate_mask_ext(const SpB_Matrix A, uint32_t *mask);
/* Exhaustive autotune: builds every applicable candidate, times it with
 * hipGraph replay (best of 3 means over reps SpMVs) and returns the times
 * (ms; INFINITY for candidates that do not apply) and the fastest. */
SpB_Info vnec_autotune_ext(const SpB_Matrix A, int reps, double times_ms[VNEC_NCAND], int *best);

/* Integer statistics of A computed on the GPU (order-independent integer
 * sums, maxima and histograms, so identical to a CPU computation): */
#define VNEC_NFEAT_INTS 32
#define VNEC_NFEAT 32
#define VNEC_FEAT_NBUCKETS 16
enum {
    VF_NROWS = 0,
    VF_NCOLS,
    VF_NNZ,
    VF_SUM_LEN2,    /* sum of squared row lengths (mod 2^64)                   */
    VF_SUM_LEN3,    /* sum of cubed row lengths (mod 2^64)                     */
    VF_MAX_LEN,     /* longest row                                              */
    VF_EMPTY_ROWS,  /* rows without entries                                     */
    VF_HIST,        /* 16 buckets: rows of length in [2^b, 2^(b+1)), b = 15 open */
    VF_GINI_NUM = VF_HIST + VNEC_FEAT_NBUCKETS,
                    /* sum over row pairs i < j of |l_i - l_j|, l = min(len, 65535) */
    VF_SUM_DIST,    /* sum over entries of |col - row * ncols / nrows|          */
    VF_MAX_DIST,    /* bandwidth: largest such distance                         */
    VF_NEAR_DIAG,   /* entries with distance <= 32                              */
    VF_SLICE_SLOTS, /* slots of the 64-row slice layout (64 x longest row)      */
    VF_NEC_SMALL,   /* distinct columns summed over merge-path windows of 4096 items */
    VF_RUNS_SMALL,  /* runs of consecutive distinct columns, same windows       */
    VF_NEC_LARGE,   /* as above, windows of 16384 items                         */
    VF_RUNS_LARGE,
};
SpB_Info vnec_feature_ints_ext(const SpB_Matrix A, uint64_t ints[VNEC_NFEAT_INTS]);
/* The selector's 32 FP32 inputs from the integer statistics. */
void vnec_features_from_ints_ext(const uint64_t ints[VNEC_NFEAT_INTS], float f[VNEC_NFEAT]);
/* Candidate mask from the integer statistics (the rules of
 * vnec_candidate_mask_ext that the statistics decide: the pattern-slice
 * bit is set whenever the slots fit, as the row pattern is not among the
 * statistics). */
uint32_t vnec_candidate_mask_from_ints_ext(const uint64_t ints[VNEC_NFEAT_INTS]);

/* MLP 32 -> 64 -> 64 -> VNEC_NCAND, ReLU, FP32.  Output c predicts
 * log2(time of candidate c / time of the fastest candidate).  Parameters
 * are one FP32 array: W1[32][64], b1[64], W2[64][64], b2[64],
 * W3[64][VNEC_NCAND], b3[VNEC_NCAND] (input-major matrices). */
#define VNEC_MLP_HIDDEN 64
#define VNEC_MLP_NPARAMS                                                                        \
    (VNEC_NFEAT * VNEC_MLP_HIDDEN + VNEC_MLP_HIDDEN + VNEC_MLP_HIDDEN * VNEC_MLP_HIDDEN +      \
     VNEC_MLP_HIDDEN + VNEC_MLP_HIDDEN * VNEC_NCAND + VNEC_NCAND)
/* The following run HIP kernels; all buffers are host memory. */
/* out[B][VNEC_NCAND] = MLP(X[B][32]) (X already standardised). */
SpB_Info vnec_mlp_forward_ext(const float *params, const float *X, int64_t B, float *out);
/* Loss sum_{b,c} mask (out - T)^2 / sum mask and its gradient (NPARAMS). */
SpB_Info vnec_mlp_grad_ext(const float *params, const float *X, const float *T, const float *mask,
                           int64_t B, float *grad, float *loss);
/* Adam training on the GPU: epochs over N samples in minibatches of
 * `batch` (order reshuffled each epoch from seed); params updated in
 * place; loss_hist (epochs entries, may be NULL) gets each epoch's mean
 * loss. */
SpB_Info vnec_mlp_train_ext(float *params, const float *X, const float *T, const float *mask,
                            int64_t N, int epochs, int batch, float lr, float weight_decay,
                            uint64_t seed, float *loss_hist);
/* Inference kernel: choice[b] = argmin over candidates allowed by mask[b]
 * of MLP((features[b] - mu) / sd) (lowest index on ties; -1 if none). */
SpB_Info vnec_mlp_select_ext(const float *params, const float *mu, const float *sd,
                             const float *features, const uint32_t *mask, int64_t B,
                             int32_t *choice);
/* Install the model VNEC_SELECT_LEARNED uses (copied; NULL params restores
 * the model compiled into the library).  *trained reports whether a model
 * is in force. */
SpB_Info vnec_selector_install_ext(const float *params, const float *mu, const float *sd);
SpB_Info vnec_selector_info_ext(int *trained, int64_t *train_samples);
/* The candidate a selection mode picks for A (single right-hand side). */
SpB_Info vnec_select_ext(const SpB_Matrix A, int mode, int *candidate);

/* ------------------------------------------------------- plan lifetimes
 * The reference has no destroy call for its plans; these release both the
 * public host fields and the device plan behind them. */
SpB_Info VNEC_D_free_fp64_ext(VNEC_D_FP64 *plan);
SpB_Info VNEC_S_free_fp64_ext(VNEC_S_FP64 *plan);
SpB_Info VNEC_D_free_fp32_ext(VNEC_D_FP32 *plan);
SpB_Info VNEC_S_free_fp32_ext(VNEC_S_FP32 *plan);
SpB_Info VNEC_L_free_ext(VNEC_L *plan);
/* Device plan behind a Build_VNEC_* result (owned by it). */
vnec_plan VNEC_plan_of_ext(const void *reference_plan);

/* ------------------------------------------- SpMV that allocates nothing
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): y_dev <- A x_dev for an FP64 CSR matrix
 * given as caller device arrays (ptr_dev: nrows + 1 int32 row offsets,
 * col_dev / val_dev: nnz int32 columns and FP64 values), enqueued on
 * `stream` (NULL: the library stream) and not waited for.  It runs the
 * kernels of vnec_spmv_dev_ext on an L plan at 7 items per thread with one
 * workgroup per tile (same tiles, same sums, bit-identical y), keeping the
 * tile coordinates and carries in a fixed device array of about 1 MiB per
 * GPU instead of plan memory: at most VNEC_NOALLOC_MAX_TILES tiles of 1792
 * merge items, i.e. nrows + nnz <= 117,440,512; larger inputs fail with a
 * message before any device work.  Calls that overlap on one GPU race on
 * that array. */
#define VNEC_NOALLOC_MAX_TILES 65536
SpB_Info vnec_csr_spmv_noalloc_ext(int64_t nrows, int64_t ncols, int64_t nnz, const int32_t *ptr_dev,
                                   const int32_t *col_dev, const double *val_dev, const double *x_dev,
                                   double *y_dev, void *stream);

#ifdef __cplusplus
}
#endif
