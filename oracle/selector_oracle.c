/*
 * selector_oracle.c — CPU restatement of the integer matrix statistics the
 * variant selector reads.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the library's GPU statistics
 * kernels (the selector has no reference counterpart; its definitions are
 * the library's own, stated in include/vnec_ext.h).  Where the GPU side
 * uses a shortcut (a length histogram for the pairwise-difference sum,
 * atomics for the sums), this file takes the direct route (sorting, plain
 * loops), so that a mistake in either shows up as a mismatch.  All sums
 * wrap modulo 2^64, like the GPU's.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int32_t x, y; } orc_coord;
void orc_partition(const int64_t *ptr, int64_t nrows, int64_t nnz, int64_t P, orc_coord *start,
                   orc_coord *end);

enum {
    F_NROWS = 0, F_NCOLS, F_NNZ, F_SUM_LEN2, F_SUM_LEN3, F_MAX_LEN, F_EMPTY, F_HIST,
    F_GINI = F_HIST + 16, F_SUM_DIST, F_MAX_DIST, F_NEAR_DIAG, F_SLICE_SLOTS,
    F_NEC_SMALL, F_RUNS_SMALL, F_NEC_LARGE, F_RUNS_LARGE, F_N
};

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Distinct columns and runs of consecutive distinct columns per tile of the
 * merge-path partition into ceil(M / items) tiles, summed. */
static void window_sums(const int64_t *ptr, const int32_t *col, int64_t nrows, int64_t nnz,
                        int64_t items, uint64_t *nec_sum, uint64_t *runs_sum)
{
    *nec_sum = *runs_sum = 0;
    if (nnz == 0) return;
    int64_t M = nrows + nnz, P = (M + items - 1) / items;
    orc_coord *s = (orc_coord *)malloc(sizeof(orc_coord) * (size_t)P);
    orc_coord *e = (orc_coord *)malloc(sizeof(orc_coord) * (size_t)P);
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (size_t)(items + 1));
    orc_partition(ptr, nrows, nnz, P, s, e);
    for (int64_t t = 0; t < P; t++) {
        int64_t n = e[t].y - s[t].y;
        memcpy(buf, col + s[t].y, sizeof(int32_t) * (size_t)n);
        qsort(buf, (size_t)n, sizeof(int32_t), cmp_i32);
        for (int64_t i = 0; i < n; i++) {
            if (i > 0 && buf[i] == buf[i - 1]) continue;
            (*nec_sum)++;
            /* a run starts where the previous distinct column is not col - 1 */
            int64_t p = i - 1;
            if (p < 0 || buf[p] != buf[i] - 1) (*runs_sum)++;
        }
    }
    free(s);
    free(e);
    free(buf);
}

void orc_feature_ints(const int64_t *ptr, const int32_t *col, int64_t nrows, int64_t ncols,
                      uint64_t *out)
{
    memset(out, 0, sizeof(uint64_t) * F_N);
    int64_t nnz = ptr[nrows];
    out[F_NROWS] = (uint64_t)nrows;
    out[F_NCOLS] = (uint64_t)ncols;
    out[F_NNZ] = (uint64_t)nnz;
    uint32_t *cl = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(nrows ? nrows : 1));
    for (int64_t r = 0; r < nrows; r++) {
        uint64_t len = (uint64_t)(ptr[r + 1] - ptr[r]);
        out[F_SUM_LEN2] += len * len;
        out[F_SUM_LEN3] += len * len * len;
        if (len > out[F_MAX_LEN]) out[F_MAX_LEN] = len;
        if (len == 0) {
            out[F_EMPTY]++;
        } else {
            int b = 0;
            while (b < 15 && (len >> (b + 1)) != 0) b++;
            out[F_HIST + b]++;
        }
        cl[r] = (uint32_t)(len < 65535 ? len : 65535);
        int64_t diag = r * ncols / nrows;
        for (int64_t j = ptr[r]; j < ptr[r + 1]; j++) {
            int64_t d = (int64_t)col[j] - diag;
            uint64_t ad = (uint64_t)(d < 0 ? -d : d);
            out[F_SUM_DIST] += ad;
            if (ad > out[F_MAX_DIST]) out[F_MAX_DIST] = ad;
            if (ad <= 32) out[F_NEAR_DIAG]++;
        }
    }
    /* sum_{i<j} |l_i - l_j| = sum_k (2k - n + 1) l_(k) over ascending order */
    qsort(cl, (size_t)nrows, sizeof(uint32_t), cmp_u32);
    for (int64_t k = 0; k < nrows; k++)
        out[F_GINI] += (uint64_t)(2 * k - nrows + 1) * (uint64_t)cl[k];
    free(cl);
    for (int64_t s0 = 0; s0 < nrows; s0 += 64) {
        uint64_t m = 0;
        for (int64_t r = s0; r < s0 + 64 && r < nrows; r++) {
            uint64_t len = (uint64_t)(ptr[r + 1] - ptr[r]);
            if (len > m) m = len;
        }
        out[F_SLICE_SLOTS] += 64 * m;
    }
    window_sums(ptr, col, nrows, nnz, 4096, &out[F_NEC_SMALL], &out[F_RUNS_SMALL]);
    window_sums(ptr, col, nrows, nnz, 16384, &out[F_NEC_LARGE], &out[F_RUNS_LARGE]);
}
