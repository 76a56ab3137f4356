/*
 * vnec_hip.h — the thin C-ABI between the C host code of libvnec_mi355x.so
 * and its hand-written HIP kernels for gfx950.  Plain pointers (device
 * addresses unless named *_host) and sizes only; every call enqueues on the
 * given stream and returns 0 on success or a hipError_t code.
 *
 * Internal to the library (no reference counterpart); declared here so the
 * boundary between host logic and kernels is explicit and testable.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Merge-path coordinates of diagonals min(t * ceil(M / P), M), t = 0..P,
 * written as (x, y) int pairs to coords (2 (P + 1) ints).  snap != 0
 * applies the S-variant row snap.  ptr: nrows + 1 int32 row offsets. */
int vnec_hip_partition(const int32_t *ptr, int64_t nrows, int64_t nnz, int64_t P, int snap,
                       int32_t *coords, void *stream);

/* Exclusive sum scan of n int32 values (out may alias in); *total_host
 * (if not NULL) receives the sum after a stream synchronisation.  Allocates
 * its own scratch (builder use only). */
int vnec_hip_scan_excl_i32(const int32_t *in, int32_t *out, int64_t n, int64_t *total_host,
                           void *stream);

/* ECR by per-tile LDS sort (tiles of at most VNEC_ECR_TILE_MAX_BIG
 * nonzeros; tiles up to VNEC_ECR_TILE_MAX use the smaller-LDS form).
 * Pass 1 writes the tile-local rank of each nonzero (lcol16 and/or ecr32,
 * either may be NULL), the sorted distinct columns of tile t at
 * uniq[y_t .. y_t + nec[t]), nec[t] and the number of runs of consecutive
 * columns nruns[t].  Pass 2 writes each tile's runs (first column, window
 * position) at run_off[t]... */
#define VNEC_ECR_TILE_MAX 4096
#define VNEC_ECR_TILE_MAX_BIG 16384
int vnec_hip_ecr_tiles_pass1(const int32_t *col, const int32_t *coords, int64_t ntiles,
                             uint16_t *lcol16, int32_t *ecr32, int32_t *uniq, int32_t *nec,
                             int32_t *nruns, void *stream);
int vnec_hip_ecr_tiles_pass2(const int32_t *coords, int64_t ntiles, const int32_t *uniq,
                             const int32_t *nec, const int32_t *run_off, int32_t *run_start,
                             int32_t *run_wpos, void *stream);
/* Largest nonzero count of any tile (device reduction, result to host). */
int vnec_hip_max_tile_nnz(const int32_t *coords, int64_t ntiles, int32_t *max_host, void *stream);

/* ECR by dense flags over P x ncols (few large partitions).  flags and rank
 * are caller scratch of P * ncols int32 each.  Writes ecr (nnz), win
 * (global ranks: win_off[t] = rank[t * ncols]) and win_off (P + 1). */
int vnec_hip_ecr_dense(const int32_t *col, int64_t nnz, int64_t ncols, const int32_t *coords,
                       int64_t P, int32_t *flags, int32_t *rank, int32_t *ecr, int32_t *win,
                       int64_t *win_off, void *stream);

/* VNEC-D block layout at (P, L) from ECR ranks ecr (nnz): pass 1 writes
 * each row's group count to cnt (nrows + 1 entries, cnt[nrows] = 0; its
 * exclusive scan is v_row_ptr) and sets *unsorted (device int, caller-
 * zeroed, may be NULL) when some row's columns descend (the greedy
 * grouping needs them ascending; pass 2 skips such entries); pass 2 writes col_start (G) and, if
 * val_align is not NULL, val_align (G * L, caller-zeroed).  A column
 * repeated within a row overwrites its slot (accumulate = 0, the
 * reference's layout) or is added to it (accumulate = 1, for SpMV). */
int vnec_hip_d_count(const int32_t *ptr, const int32_t *ecr, int64_t nrows, const int32_t *coords,
                     int64_t P, int L, int32_t *cnt, int32_t *unsorted, void *stream);
int vnec_hip_d_fill(const int32_t *ptr, const int32_t *ecr, const double *val, int64_t nrows,
                    const int32_t *coords, int64_t P, int L, const int32_t *v_row_ptr,
                    int32_t *col_start, double *val_align, int accumulate, void *stream);

/* Right-hand-side layouts, k <= 16 vectors of n doubles: column-major
 * (src[c n + j]) to interleaved (dst[j k + c]) when to_interleaved, else
 * interleaved to column-major. */
int vnec_hip_rhs_layout(const double *src, int64_t n, int k, int to_interleaved, double *dst, void *stream);

/* FP64 -> FP32 hi + FP32 lo split: hi = fl32(v), lo = fl32(v - hi). */
int vnec_hip_split_hilo(const double *val, int64_t n, float *hi, float *lo, void *stream);

/* Merge-path SpMV arguments (variant L or E; see vnec_ext.h). */
typedef struct {
    const int32_t *ptr;       /* nrows + 1                         */
    const int32_t *col;       /* nnz, global column (variant L)    */
    const uint16_t *lcol;     /* nnz, tile-local rank (variant E)  */
    const double *val;        /* nnz, encoding FP64                */
    const float *hi, *lo;     /* nnz, encodings HILO / HI          */
    const int32_t *coords;    /* 2 (ntiles + 1)                    */
    int64_t ntiles;
    const int32_t *run_off;   /* ntiles + 1 (variant E)            */
    const int32_t *run_start; /* runs (variant E)                  */
    const int32_t *run_wpos;  /* runs (variant E)                  */
    const int32_t *nec;       /* ntiles (variant E)                */
    const double *x;          /* ncols * nrhs                      */
    double *y;                /* nrows * nrhs                      */
    double *carry;            /* ntiles * nrhs scratch             */
    int64_t nrows;
    int ipt;
    int variant;
    int encoding;
    int nrhs;
    const double *xhot;       /* nhot packed hot x entries (variant H) */
    int nhot;
    /* encoding FP32 (FP32 values in hi, FP32 x / y / sums) */
    const float *x32;         /* ncols                             */
    float *y32;               /* nrows                             */
    float *carry32;           /* ntiles scratch                    */
    /* variant C: val / col hold each tile's nonzeros in column order and
       spos their position within the tile's CSR range; with cpos_shift > 0
       the position is packed into col instead: col = column | pos << shift,
       the column taken relative to cbase[t] when cbase is given */
    const uint16_t *spos;     /* nnz                               */
    int cpos_shift;
    const int32_t *cbase;     /* ntiles, or NULL                   */
    int block;                /* threads per tile of the persistent
                                 kernel: 0 (256) or VNEC_WIDE_BLOCK */
    int gather;               /* strided launch of variant C: 1 when each
                                 thread issues all of a tile's x gathers
                                 before using any                  */
} vnec_merge_args;

/* Threads per tile of the wide column-sorted tiles (variant C, FP64, one
 * right-hand side, persistent launch; ipt 3, 5, 7, 9 or 11): a tile holds
 * VNEC_WIDE_BLOCK * ipt items and its nonzeros are sorted by column over
 * that whole window. */
#define VNEC_WIDE_BLOCK 512

int vnec_hip_spmv_merge(const vnec_merge_args *a, void *stream);
/* Adds each tile's carry to y[row] in tile order (one add per row; rows
 * >= nrows, the carries of empty trailing tiles, are skipped). */
int vnec_hip_carry_fixup(const int32_t *coords, const double *carry, int64_t ntiles, double *y,
                         int64_t nrows, int nrhs, void *stream);
/* The same for FP32 carries and y (one right-hand side). */
int vnec_hip_carry_fixup_f32(const int32_t *coords, const float *carry, int64_t ntiles, float *y,
                             int64_t nrows, void *stream);
/* Persistent pipelined form (one right-hand side): nwg workgroups (a
 * multiple of 8) each own a contiguous run of tiles; each leaves the
 * partial row at its end in wg_row / wg_carry (nwg entries) for
 * vnec_hip_wg_fixup.  vnec_hip_stream_grid gives the resident workgroup
 * count (occupancy x CUs, rounded down to a multiple of 8). */
int vnec_hip_stream_grid(int variant, int ipt, int encoding, int block, int *nwg);
/* Whether the persistent kernel exists for (variant, ipt, encoding, block);
 * block 0 is 256 threads per tile, VNEC_WIDE_BLOCK the wide tiles. */
int vnec_hip_stream_supported(int variant, int ipt, int encoding, int block);
/* The pipelined variant-E kernel holds at most this many window runs per
 * tile (one per thread); plans with a larger tile fall back to one
 * workgroup per tile. */
#define VNEC_E_STREAM_MAX_RUNS 256
int vnec_hip_spmv_stream(const vnec_merge_args *a, int nwg, int32_t *wg_row, double *wg_carry,
                         void *stream);
int vnec_hip_wg_fixup(const int32_t *wg_row, const double *wg_carry, int nwg, double *y,
                      int64_t nrows, void *stream);
/* Strided persistent form of the one-workgroup-per-tile kernel (variant L
 * in every encoding, variant C in FP64 and the FP32 API's encoding, also in
 * wide tiles, and for C in FP64 at ipt 3, 5 or 7 with each thread's gathers
 * issued together when gather = 1 (a.gather); one right-hand side): the
 * same per-tile carries in a.carry
 * (for vnec_hip_carry_fixup) and the same results bit for bit; nwg
 * workgroups (a multiple of 8), those on XCD x sweeping x's share of the
 * tiles together.  vnec_hip_strided_grid gives the resident count;
 * vnec_hip_strided_supported whether the kernel exists. */
int vnec_hip_strided_supported(int variant, int ipt, int encoding, int block, int gather);
int vnec_hip_strided_grid(int variant, int ipt, int encoding, int block, int gather, int *nwg);
int vnec_hip_spmv_strided(const vnec_merge_args *a, int nwg, void *stream);
/* Variant H: the persistent kernel with the hot x entries in LDS (hot =
 * 4096 or 8192 slots at 5 or 7 items per thread; FP64, one right-hand
 * side).  col holds -1 - slot for hot columns.  Workgroups of
 * 1024 threads run four tile pipelines over one LDS table: nwg workgroups
 * own 4 nwg runs of tiles, and wg_row / wg_carry hold 4 nwg entries. */
int vnec_hip_hot_supported(int ipt, int hot);
int vnec_hip_hot_grid(int ipt, int hot, int *nwg);
int vnec_hip_spmv_hot(const vnec_merge_args *a, int hot, int nwg, int32_t *wg_row, double *wg_carry,
                      void *stream);
/* Hot-column set of a matrix: deg (ncols, caller-zeroed) gets each column's
 * use count; vnec_hip_hot_order sorts columns by descending count (ties by
 * column) into order (ncols); vnec_hip_hot_encode writes slot[] (ncols, -1
 * for cold columns) from order[0 .. nhot) and the encoded columns hcol
 * (nnz). */
int vnec_hip_col_degree(const int32_t *col, int64_t nnz, int32_t *deg, void *stream);
int vnec_hip_hot_order(const int32_t *deg, int64_t ncols, int32_t *order, void *stream);
int vnec_hip_hot_encode(const int32_t *col, int64_t nnz, const int32_t *order, int64_t nhot,
                        int64_t ncols, int32_t *slot, int32_t *hcol, void *stream);
/* order (ncols): the first k columns of order_full, then the others in
 * ascending column order (rank, flags: ncols scratch each). */
int vnec_hip_pack_order(const int32_t *order_full, int64_t ncols, int64_t k, int32_t *rank, int32_t *flags,
                        int32_t *order, void *stream);
/* Relabelled columns: rank[order[h]] = h (rank: ncols scratch) and
 * rcol[j] = rank[col[j]] (nnz), order a permutation of the columns. */
int vnec_hip_relabel_cols(const int32_t *col, int64_t nnz, const int32_t *order, int64_t ncols,
                          int32_t *rank, int32_t *rcol, void *stream);
/* xhot[h] = x[order[h]], h < nhot (before each variant-H SpMV, and with
 * nhot = ncols the x permutation of a relabelled plan). */
int vnec_hip_gather_hot(const double *x, const int32_t *order, int64_t nhot, double *xhot,
                        void *stream);
/* Variant C tile storage: per tile t (nonzeros coords[t].y .. coords[t+1].y)
 * a stable sort by column; sval / scol / spos (nnz) receive the values,
 * columns and in-tile positions in that order.  With pos_shift > 0 the
 * positions go into scol's high bits (scol = column | pos << pos_shift,
 * columns below 2^pos_shift, positions below 2^(32 - pos_shift)) and spos
 * is not written; with cbase (ntiles, from vnec_hip_tile_colspan) the packed
 * column is column - cbase[t], which must stay below 2^pos_shift.  keys /
 * idx: nnz scratch. */
int vnec_hip_csort(const int32_t *coords, int64_t ntiles, const int32_t *col, const double *val,
                   int64_t nnz, int64_t ncols, int pos_shift, const int32_t *cbase, uint64_t *keys,
                   int32_t *idx, double *sval, int32_t *scol, uint16_t *spos, void *stream);
/* cmin[t] = the smallest column of tile t (0 for a tile with no nonzeros);
 * *span_host = the largest cmax - cmin + 1 over the tiles (after a stream
 * synchronisation). */
int vnec_hip_tile_colspan(const int32_t *coords, int64_t ntiles, const int32_t *col, int32_t *cmin,
                          int32_t *span_host, void *stream);
/* Whether a (variant, ipt, encoding, nrhs) combination is compiled in. */
int vnec_hip_merge_supported(int variant, int ipt, int encoding, int nrhs);

/* Gather-ceiling probe over a matrix's values and columns (nnz): with
 * gather = 0 it streams them, with gather = 1 it also gathers x[col] per
 * nonzero; no row structure; out holds grid * 256 partial sums. */
int vnec_hip_gather_probe(const double *val, const int32_t *col, int64_t nnz, const double *x,
                          int gather, double *out, int grid, void *stream);

/* Streaming copy of bytes (multiple of 16) from src to dst. */
int vnec_hip_stream_copy(const void *src, void *dst, size_t bytes, void *stream);
/* Streaming read of bytes (a multiple of 8) from src, one contiguous 64 KiB
 * chunk per workgroup; out (>= 4096 words) is written only for an input
 * whose words XOR-fold to 0x9E3779B9 per thread (a probe or a cache flush:
 * pass bytes of a memset pattern). */
int vnec_hip_stream_read(const void *src, size_t bytes, void *out, void *stream);
/* The same read with cache-allocating loads: reading more than the L2 and
 * Infinity Cache hold evicts their contents without dirtying them. */
int vnec_hip_flush_read(const void *src, size_t bytes, void *out, void *stream);

/* ---- variant S: 64-row slices, [l][lane] layout ---- */
typedef struct {
    const int32_t *ptr;         /* nrows + 1 (row lengths)                 */
    const int32_t *slice_len;   /* nslices: longest row of each slice      */
    const int32_t *slice_ptr;   /* nslices + 1: padded offsets              */
    const double *sv;           /* padded values, encoding FP64             */
    const float *shi, *slo;     /* padded values, encodings HILO / HI       */
    const int32_t *scol;        /* padded global columns (win == 0)         */
    const uint16_t *slcol;      /* padded window ranks (win == 1)           */
    const int32_t *tile_slices; /* 2 (ntiles + 1): (first slice, offset)    */
    int64_t ntiles;
    const int32_t *run_off, *run_start, *run_wpos, *nec; /* windows        */
    int32_t max_nec;            /* LDS window entries per workgroup         */
    int win;
    const double *x;
    double *y;
    int64_t nrows;
    int encoding;
    const float *x32;           /* encoding FP32: x, y in FP32, values shi  */
    float *y32;
} vnec_slice_args;

/* slice_len[s] = longest row of slice s, slice_sz[s] = 64 * slice_len[s],
 * slice_sz[nslices] = 0 (its exclusive scan gives the padded offsets). */
int vnec_hip_slice_len(const int32_t *ptr, int64_t nrows, int64_t nslices, int32_t *slice_len,
                       int32_t *slice_sz, void *stream);
/* (row, nonzero) coordinates of tile boundaries given in slice space. */
int vnec_hip_slice_tile_rows(const int32_t *slice_coords, int64_t ntiles, const int32_t *ptr,
                             int64_t nrows, int32_t *row_coords, void *stream);
/* Transposed, padded slice arrays (any output may be NULL). */
int vnec_hip_slice_fill(const int32_t *ptr, const int32_t *col, const uint16_t *lcol,
                        const double *val, int64_t nrows, int64_t nslices,
                        const int32_t *slice_len, const int32_t *slice_ptr, double *sv, float *shi,
                        float *slo, int32_t *scol, uint16_t *slcol, void *stream);
int vnec_hip_spmv_slice(const vnec_slice_args *a, void *stream);

/* ---- variant D: dense column blocks of the ECR window ---- */
typedef struct {
    const int32_t *ptr;       /* nrows + 1                                     */
    const int32_t *vrp;       /* nrows + 1: groups of row i are [vrp[i], vrp[i+1]) */
    const int32_t *col_start; /* groups: first window position of each group   */
    const double *vals;       /* groups * L, blocks of 64 groups stored [j][lane] */
    const double *val;        /* nnz: CSR values (each tile's carried row)     */
    const uint16_t *lcol;     /* nnz: tile-local window ranks                  */
    const int32_t *coords;    /* 2 (ntiles + 1)                                */
    int64_t ntiles;
    const int32_t *run_off, *run_start, *run_wpos, *nec; /* windows            */
    const double *x;
    double *y;
    double *carry;            /* ntiles                                        */
    int64_t nrows;
    int ipt;
    int lanes;                /* L: 4, 8 or 16                                 */
    int fp32;                 /* FP32 arithmetic: the *32 fields below are used */
    const float *vals32, *val32, *x32;
    float *y32, *carry32;
} vnec_d_args;

/* Wave64 layout of a D value array: in[g * L + j] -> out[((g / 64) * L + j)
 * * 64 + g % 64]; out holds ceil(G / 64) * 64 * L doubles. */
int vnec_hip_d_interleave(const double *in, int64_t G, int L, double *out, void *stream);
/* Same layout narrowed to FP32 (exact for values that came from FP32). */
int vnec_hip_d_interleave_f32(const double *in, int64_t G, int L, float *out, void *stream);
int vnec_hip_spmv_d(const vnec_d_args *a, void *stream);
int vnec_hip_d_supported(int ipt, int lanes);

/* ---- variant G: pattern slices ---- */
/* Longest row pattern a plan stores (positions per row; one mask bit each). */
#define VNEC_PATTERN_MAX 32
typedef struct {
    const uint32_t *masks; /* nrows: pattern positions each row holds (0: none,
                              or a row left to the sub-plan)                 */
    const double *gv;      /* nslices * L * 64 values, slice s stored [l][lane],
                              rows 64 s .. 64 s + 63, 0 where a bit is clear
                              (encoding 0)                                     */
    const float *gh, *gl;  /* the same slots as FP32 hi and lo (encoding 1), or
                              hi only (encoding 2); summed in FP64             */
    int encoding;          /* 0 FP64, 1 FP32 hi + lo, 2 FP32 hi (FP64 sums);
                              3 the FP32 API: FP32 values in gh, FP32 x32,
                              y32 and sums                                    */
    const int32_t *pat;    /* L column offsets: position l of row i is column
                              i + pat[l]                                      */
    int L;
    int64_t nrows;
    int64_t nslices;       /* ceil(nrows / 64)                                */
    const double *x;
    double *y;
    /* x window of a workgroup's VNEC_PATTERN_ROWS rows r0 ..: the union of the
       column ranges r0 + pat[l] + [0, VNEC_PATTERN_ROWS), as nwin runs (offset wst[k] from
       r0, length wlen[k], LDS start wbase[k]), staged in LDS once; position
       l of local row t reads LDS entry wpos[l] + t.  wsize = 0: direct x
       reads.  ncols bounds the staged columns. */
    int nwin, wsize;
    int32_t wst[VNEC_PATTERN_MAX], wlen[VNEC_PATTERN_MAX], wbase[VNEC_PATTERN_MAX], wpos[VNEC_PATTERN_MAX];
    int64_t ncols;
    const float *x32;      /* encoding 3                                      */
    float *y32;
} vnec_pattern_args;
/* Largest x window staged in LDS (doubles). */
#define VNEC_PATTERN_WIN_MAX 4096
/* Rows per workgroup of the pattern kernel (the window covers their runs). */
#define VNEC_PATTERN_ROWS 512
/* masks[i] = positions of the pattern row i holds; flags[i] = 1 when row i
 * is not a subset of the pattern (a column outside it, or repeated). */
int vnec_hip_pattern_match(const int32_t *ptr, const int32_t *col, int64_t nrows, const int32_t *pat,
                           int L, uint32_t *masks, int32_t *flags, void *stream);
/* With pos = exclusive scan of flags: irr[pos[i]] = i for flagged rows. */
int vnec_hip_pattern_split(const int32_t *flags, const int32_t *pos, int64_t nrows, int32_t *irr,
                           void *stream);
/* gv[(s L + l) 64 + lane] = value of row 64 s + lane at position l (0 when
 * the row does not hold it, or past nrows). */
int vnec_hip_pattern_fill(const int32_t *ptr, const double *val, const uint32_t *masks, int64_t nrows,
                          int L, double *gv, void *stream);
/* y[i] = sum over held positions l of value * x[i + pat[l]], every row i;
 * values from gv, or gh (+ gl) by a->encoding. */
int vnec_hip_spmv_pattern(const vnec_pattern_args *a, void *stream);
/* y[rows[i]] = ysub[i], i < n. */
int vnec_hip_scatter_rows(const double *ysub, const int32_t *rows, int64_t n, double *y, void *stream);
int vnec_hip_scatter_rows_f32(const float *ysub, const int32_t *rows, int64_t n, float *y, void *stream);

/* ---- VNEC-S task order and reordered matrix (builder) ---- */
/* Stable sort of n (key, value) pairs by the low key_bits bits of the keys
 * (LSD radix, 8-bit digits); allocates its own scratch. */
int vnec_hip_sort_pairs_u64(uint64_t *keys, int32_t *vals, int64_t n, int key_bits, void *stream);
/* Task list (nrows) and cross-row counts (P) of the snapped partitions in
 * coords (2 (P + 1) ints): per partition, rows by length then row id, the
 * first c - c % L rows of each length <= cut first. */
int vnec_hip_s_order(const int32_t *ptr, int64_t nrows, const int32_t *coords, int64_t P, int L,
                     int cut, int32_t *tasks, int32_t *spvv_len, void *stream);
/* Rows copied in task order (row_begin / row_end: nrows each; s_val may be
 * NULL; s_col gets the ECR ranks ecr[]), cross-row groups as [l][row]. */
int vnec_hip_s_reorder(const int32_t *ptr, const double *val, const int32_t *ecr,
                       const int32_t *tasks, int64_t nrows, int64_t nnz, const int32_t *coords,
                       const int32_t *spvv_len, int64_t P, int L, int32_t *row_begin,
                       int32_t *row_end, double *s_val, int32_t *s_col, void *stream);

/* y[0 .. n) = 0 (a kernel, safe inside captured graphs). */
int vnec_hip_zero_f64(double *y, int64_t n, void *stream);
int vnec_hip_zero_f32(float *y, int64_t n, void *stream);
/* out[i] = (double)in[i], i < n (exact). */
int vnec_hip_widen_f32(const float *in, int64_t n, double *out, void *stream);

/* Overwrites n bytes (cache flush helper for timing). */
int vnec_hip_scribble(void *p, size_t n, void *stream);

/* ---- selector: integer statistics (stats: 32 uint64, caller-zeroed;
 * lenhist: 65536 uint32 counts of min(row length, 65535), caller-zeroed) */
int vnec_hip_feature_rows(const int32_t *ptr, int64_t nrows, uint64_t *stats, uint32_t *lenhist,
                          void *stream);
int vnec_hip_feature_cols(const int32_t *ptr, const int32_t *col, int64_t nrows, int64_t ncols,
                          uint64_t *stats, void *stream);

/* ---- selector: MLP kernels (FP32 parameters laid out as in vnec_ext.h) */
int vnec_hip_mlp_forward(const float *params, const float *X, int64_t B, float *out, void *stream);
/* Gradient partials of the masked squared loss scaled by inv_den over the
 * samples idx[0..B) (idx may be NULL: samples 0..B-1), chunks of `chunk`
 * samples: part[ceil(B / chunk)][NPARAMS], loss_part[ceil(B / chunk)]. */
int vnec_hip_mlp_grad(const float *params, const float *X, const float *T, const float *mask,
                      const int32_t *idx, int64_t B, int chunk, float inv_den, float *part,
                      float *loss_part, void *stream);
/* Sums nchunks partials in chunk order into grad (may be NULL) and, if
 * lr > 0, applies one Adam step (bias corrections bc1 = 1 - beta1^t,
 * bc2 = 1 - beta2^t) to params with state m, v. */
int vnec_hip_mlp_reduce_adam(const float *part, int nchunks, float *grad, float *params, float *m,
                             float *v, float lr, float weight_decay, float bc1, float bc2,
                             void *stream);
int vnec_hip_mlp_select(const float *params, const float *mu, const float *sd,
                        const float *features, const uint32_t *mask, int64_t B, int32_t *choice,
                        void *stream);

#ifdef __cplusplus
}
#endif
