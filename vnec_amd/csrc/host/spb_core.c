/*
 * spb_core.c — library runtime and the OpenSparseBLAS object subset of the
 * VNEC path (CSR matrix and dense vector, FP64 or FP32), with device
 * mirrors.
 *
 * Behaviour follows the reference objects (src/core/src/SpB_Matrix.c,
 * SpB_Matrix_CSC_or_CSR.c, SpB_Vector.c, SpB_Vector_Dense.c): zero-filled
 * vectors, row-major CSR with 64-bit host indices, duplicates kept by
 * build with SpB_BINARYOP_NULL.  Errors return SpB_FAILURE (never exit).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spb_internal.h"

/* ---------------------------------------------------------------- errors */
static __thread char g_err[512];

const char *vnec_last_error(void) { return g_err; }

void vnec_clear_error(void) { g_err[0] = '\0'; }

SpB_Info vnec_fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return SpB_FAILURE;
}

SpB_Info vnec_hip_ok(int rc, const char *what)
{
    if (rc == 0) return SpB_SUCCESS;
    return vnec_fail("%s: HIP error %d (%s)", what, rc, hipGetErrorString((hipError_t)rc));
}

/* --------------------------------------------------------------- runtime */
static hipStream_t g_stream = NULL;
static int g_rt_state = 0; /* 0 unknown, 1 ready, -1 no device */
static int g_csc_default = 1; /* the reference's default format is CSC (SpB_Global.c:17) */

int vnec_runtime_ready(void)
{
    if (g_rt_state == 0) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
            g_rt_state = -1;
        } else if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) {
            g_rt_state = -1;
        } else {
            g_rt_state = 1;
        }
    }
    if (g_rt_state < 0) vnec_fail("no HIP device available");
    return g_rt_state > 0;
}

hipStream_t vnec_stream(void) { return g_stream; }

int vnec_device_count_ext(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void *vnec_malloc_dev(size_t bytes, int64_t *acct)
{
    void *p = NULL;
    if (bytes == 0) bytes = 16;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        vnec_fail("hipMalloc(%zu) failed", bytes);
        return NULL;
    }
    if (acct) *acct += (int64_t)bytes;
    return p;
}

void vnec_free_dev(void *p)
{
    if (p) hipFree(p);
}

void SpB_init()
{
    g_csc_default = 1;
    (void)vnec_runtime_ready();
}

void SpB_finalize()
{
    if (g_rt_state > 0) {
        hipStreamSynchronize(g_stream);
        hipStreamDestroy(g_stream);
        g_stream = NULL;
        g_rt_state = 0;
    }
}

SpB_Info SpB_Global_Option_set_ext(SpB_Option_Field_ext field, SpB_Object_Format_ext value)
{
    switch (field) {
    case SpB_MATRIX_FORMAT:
        if (value.matrix_format != SpB_CSR && value.matrix_format != SpB_CSC)
            return vnec_fail("unknown matrix format %d", (int)value.matrix_format);
        g_csc_default = value.matrix_format == SpB_CSC;
        return SpB_SUCCESS;
    case SpB_VECTOR_FORMAT:
        if (value.vector_format != SpB_DENSE)
            return vnec_fail("only dense vectors are on the VNEC path");
        return SpB_SUCCESS;
    default:
        return vnec_fail("unknown global option %d", (int)field);
    }
}

/* ---------------------------------------------------------------- matrix */
spb_csr *spb_as_csr(const SpB_Matrix A)
{
    spb_csr *a = (spb_csr *)A;
    if (!a || a->magic != SPB_MAGIC_MATRIX) {
        vnec_fail("not a matrix made by SpB_Matrix_new");
        return NULL;
    }
    if (a->base.format_type != SpB_CSR) {
        vnec_fail("the VNEC path needs a CSR matrix (set SpB_MATRIX_FORMAT to SpB_CSR)");
        return NULL;
    }
    return a;
}

static void *aligned64(size_t bytes)
{
    void *p = NULL;
    if (bytes == 0) bytes = 64;
    bytes = (bytes + 63) / 64 * 64;
    if (posix_memalign(&p, 64, bytes) != 0) return NULL;
    return p;
}

SpB_Info SpB_Matrix_new(SpB_Matrix *m, SpB_Type type, SpB_Index nrows, SpB_Index ncols)
{
    if (!m) return vnec_fail("SpB_Matrix_new: NULL handle");
    if (type != SpB_FP64 && type != SpB_FP32)
        return vnec_fail("SpB_Matrix_new: the VNEC path takes SpB_FP64 or SpB_FP32 values");
    if (nrows >= (SpB_Index)INT32_MAX || ncols >= (SpB_Index)INT32_MAX)
        return vnec_fail("SpB_Matrix_new: dimensions must be below 2^31 - 1");
    spb_csr *a = (spb_csr *)calloc(1, sizeof(spb_csr));
    if (!a) return vnec_fail("SpB_Matrix_new: out of memory");
    a->base.element_type = type;
    a->base.format_type = g_csc_default ? SpB_CSC : SpB_CSR;
    a->base.row = nrows;
    a->base.col = ncols;
    a->base.element_size = type == SpB_FP32 ? sizeof(float) : sizeof(double);
    a->ptr_len = nrows + 1;
    a->ptr = (SpB_Index *)calloc((size_t)nrows + 1, sizeof(SpB_Index));
    a->magic = SPB_MAGIC_MATRIX;
    a->tuned_cand = -1;
    a->tuned_fp32 = -1;
    if (!a->ptr) {
        free(a);
        return vnec_fail("SpB_Matrix_new: out of memory");
    }
    *m = (SpB_Matrix)a;
    return SpB_SUCCESS;
}

static void matrix_drop_device(spb_csr *a)
{
    if (a->auto_plan) {
        struct vnec_plan_s *p = a->auto_plan;
        a->auto_plan = NULL;
        vnec_plan_free_ext(&p);
    }
    vnec_free_dev(a->d_ptr);
    vnec_free_dev(a->d_col);
    vnec_free_dev(a->d_val);
    a->d_ptr = NULL;
    a->d_col = NULL;
    a->d_val = NULL;
    a->dev_ready = 0;
    a->tuned_cand = -1;
    a->tuned_fp32 = -1;
    a->refined_ipt = 0;
    a->feat_ready = 0;
    a->pattern_ready = 0;
    a->refined_slice_ipt = 0;
}

SpB_Info SpB_Matrix_free(SpB_Matrix *A)
{
    if (!A || !*A) return vnec_fail("SpB_Matrix_free: NULL handle");
    spb_csr *a = (spb_csr *)*A;
    if (a->magic != SPB_MAGIC_MATRIX) return vnec_fail("SpB_Matrix_free: not a matrix");
    matrix_drop_device(a);
    free(a->ptr);
    free(a->indices);
    free(a->val);
    a->magic = 0;
    free(a);
    *A = NULL;
    return SpB_SUCCESS;
}

SpB_Info SpB_Matrix_nrows(SpB_Index *n, SpB_Matrix m)
{
    if (!n || !m) return vnec_fail("SpB_Matrix_nrows: NULL argument");
    *n = m->row;
    return SpB_SUCCESS;
}

SpB_Info SpB_Matrix_ncols(SpB_Index *n, SpB_Matrix m)
{
    if (!n || !m) return vnec_fail("SpB_Matrix_ncols: NULL argument");
    *n = m->col;
    return SpB_SUCCESS;
}

SpB_Info SpB_Matrix_nvals_ext(SpB_Index *nvals, const SpB_Matrix A)
{
    spb_csr *a = (spb_csr *)A;
    if (!nvals || !a || a->magic != SPB_MAGIC_MATRIX) return vnec_fail("SpB_Matrix_nvals_ext");
    *nvals = a->nnz;
    return SpB_SUCCESS;
}

static SpB_Info matrix_reserve(spb_csr *a, SpB_Index n)
{
    matrix_drop_device(a);
    free(a->indices);
    free(a->val);
    a->indices = (SpB_Index *)aligned64(sizeof(SpB_Index) * (size_t)n);
    a->val = aligned64(a->base.element_size * (size_t)n);
    if (!a->indices || !a->val) return vnec_fail("matrix build: out of memory (%lu entries)", n);
    a->capacity = n;
    a->nnz = n;
    return SpB_SUCCESS;
}

/* Build from (I, J, X) tuples — SpB_Matrix_CSC_or_CSR.c:188-256: entries
 * sorted by (row, col), duplicates kept in input order.  A stable counting
 * sort by row then an insertion/merge sort by column per row replaces the
 * reference's qsort; every row pointer is filled. */
static int cmp_entry(const void *p, const void *q)
{
    const SpB_Index *a = (const SpB_Index *)p, *b = (const SpB_Index *)q;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    return a[1] < b[1] ? -1 : (a[1] > b[1] ? 1 : 0); /* a[1] = input position */
}

static SpB_Info matrix_build(const char *fn, SpB_Type type, SpB_Matrix C, const SpB_Index *I,
                             const SpB_Index *J, const void *X, SpB_Index nvals, const SpB_BinaryOp dup)
{
    spb_csr *a = (spb_csr *)C;
    if (!a || a->magic != SPB_MAGIC_MATRIX) return vnec_fail("%s: bad matrix", fn);
    if (a->base.element_type != type)
        return vnec_fail("%s: the matrix was made with another value type", fn);
    if (dup != SpB_BINARYOP_NULL)
        return vnec_fail("%s: only SpB_BINARYOP_NULL is supported", fn);
    if (a->base.format_type != SpB_CSR)
        return vnec_fail("%s: the VNEC path builds CSR matrices", fn);
    if (nvals >= (SpB_Index)INT32_MAX) return vnec_fail("%s: nnz >= 2^31", fn);
    if (nvals && (!I || !J || !X)) return vnec_fail("%s: NULL arrays", fn);
    SpB_Index nr = a->base.row, nc = a->base.col;
    for (SpB_Index k = 0; k < nvals; k++)
        if (I[k] >= nr || J[k] >= nc)
            return vnec_fail("%s: entry %lu (%lu, %lu) out of range", fn, k, I[k], J[k]);
    if (matrix_reserve(a, nvals) != SpB_SUCCESS) return SpB_FAILURE;
    SpB_Index *cnt = (SpB_Index *)calloc((size_t)nr + 1, sizeof(SpB_Index));
    SpB_Index *order = (SpB_Index *)malloc(sizeof(SpB_Index) * (size_t)(nvals ? nvals : 1));
    SpB_Index *pairs = (SpB_Index *)malloc(sizeof(SpB_Index) * 2 * 64);
    if (!cnt || !order || !pairs) {
        free(cnt); free(order); free(pairs);
        return vnec_fail("%s: out of memory", fn);
    }
    for (SpB_Index k = 0; k < nvals; k++) cnt[I[k] + 1]++;
    for (SpB_Index r = 0; r < nr; r++) cnt[r + 1] += cnt[r];
    for (SpB_Index r = 0; r <= nr; r++) a->ptr[r] = cnt[r];
    for (SpB_Index k = 0; k < nvals; k++) order[cnt[I[k]]++] = k; /* stable by row */
    size_t pcap = 64;
    for (SpB_Index r = 0; r < nr; r++) {
        SpB_Index b = a->ptr[r], e = a->ptr[r + 1], len = e - b;
        if (len > pcap) {
            pcap = len;
            SpB_Index *np = (SpB_Index *)realloc(pairs, sizeof(SpB_Index) * 2 * pcap);
            if (!np) {
                free(cnt); free(order); free(pairs);
                return vnec_fail("%s: out of memory", fn);
            }
            pairs = np;
        }
        for (SpB_Index k = 0; k < len; k++) {
            pairs[2 * k] = J[order[b + k]];
            pairs[2 * k + 1] = order[b + k];
        }
        qsort(pairs, (size_t)len, 2 * sizeof(SpB_Index), cmp_entry);
        for (SpB_Index k = 0; k < len; k++) {
            a->indices[b + k] = pairs[2 * k];
            if (type == SpB_FP32)
                ((float *)a->val)[b + k] = ((const float *)X)[pairs[2 * k + 1]];
            else
                ((double *)a->val)[b + k] = ((const double *)X)[pairs[2 * k + 1]];
        }
    }
    free(cnt);
    free(order);
    free(pairs);
    return SpB_SUCCESS;
}

SpB_Info SpB_Matrix_build_FP64(SpB_Matrix C, const SpB_Index *I, const SpB_Index *J, const double *X,
                               SpB_Index nvals, const SpB_BinaryOp dup)
{
    return matrix_build("SpB_Matrix_build_FP64", SpB_FP64, C, I, J, X, nvals, dup);
}

SpB_Info SpB_Matrix_build_FP32(SpB_Matrix C, const SpB_Index *I, const SpB_Index *J, const float *X,
                               SpB_Index nvals, const SpB_BinaryOp dup)
{
    return matrix_build("SpB_Matrix_build_FP32", SpB_FP32, C, I, J, X, nvals, dup);
}

static SpB_Info import_csr(SpB_Matrix A, SpB_Type type, const int64_t *ptr, const int32_t *col,
                           const void *val)
{
    spb_csr *a = (spb_csr *)A;
    if (!a || a->magic != SPB_MAGIC_MATRIX) return vnec_fail("SpB_Matrix_import_csr_ext: bad matrix");
    if (a->base.element_type != type)
        return vnec_fail("SpB_Matrix_import_csr_ext: the matrix was made with another value type");
    if (a->base.format_type != SpB_CSR)
        return vnec_fail("SpB_Matrix_import_csr_ext: the VNEC path needs CSR");
    SpB_Index nr = a->base.row, nc = a->base.col;
    if (!ptr || ptr[0] != 0) return vnec_fail("SpB_Matrix_import_csr_ext: ptr[0] must be 0");
    int64_t nnz = ptr[nr];
    if (nnz < 0 || nnz >= INT32_MAX) return vnec_fail("SpB_Matrix_import_csr_ext: nnz out of range");
    for (SpB_Index r = 0; r < nr; r++)
        if (ptr[r + 1] < ptr[r]) return vnec_fail("SpB_Matrix_import_csr_ext: ptr not monotone");
    if (matrix_reserve(a, (SpB_Index)nnz) != SpB_SUCCESS) return SpB_FAILURE;
    for (int64_t j = 0; col && j < nnz; j++) {
        if (col[j] < 0 || (SpB_Index)col[j] >= nc) {
            a->nnz = 0;
            return vnec_fail("SpB_Matrix_import_csr_ext: column %d out of range", col[j]);
        }
        a->indices[j] = (SpB_Index)col[j];
    }
    if (nnz) memcpy(a->val, val, a->base.element_size * (size_t)nnz);
    for (SpB_Index r = 0; r <= nr; r++) a->ptr[r] = (SpB_Index)ptr[r];
    return SpB_SUCCESS;
}

SpB_Info SpB_Matrix_import_csr_ext(SpB_Matrix A, const int64_t *ptr, const int32_t *col,
                                   const double *val)
{
    return import_csr(A, SpB_FP64, ptr, col, val);
}

SpB_Info SpB_Matrix_import_csr_fp32_ext(SpB_Matrix A, const int64_t *ptr, const int32_t *col,
                                        const float *val)
{
    return import_csr(A, SpB_FP32, ptr, col, val);
}

/* Block-diagonal batch of nmat matrices as one CSR matrix: member m owns
 * rows [row_off[m], row_off[m+1]) and columns [col_off[m], col_off[m+1]);
 * `col` holds member-local columns, `ptr` global row offsets.  One SpMV
 * over the batch is then one merge-path launch over all the nonzeros. */
SpB_Info vnec_batch_import_ext(SpB_Matrix A, int64_t nmat, const int64_t *row_off,
                               const int64_t *col_off, const int64_t *ptr, const int32_t *col,
                               const double *val)
{
    spb_csr *a = (spb_csr *)A;
    if (!a || a->magic != SPB_MAGIC_MATRIX || nmat < 1 || !row_off || !col_off || !ptr)
        return vnec_fail("vnec_batch_import_ext: bad arguments");
    if ((SpB_Index)row_off[nmat] != a->base.row || (SpB_Index)col_off[nmat] != a->base.col)
        return vnec_fail("vnec_batch_import_ext: matrix is %lux%lu, batch is %ldx%ld", a->base.row,
                         a->base.col, (long)row_off[nmat], (long)col_off[nmat]);
    if (row_off[0] != 0 || col_off[0] != 0) return vnec_fail("vnec_batch_import_ext: offsets must start at 0");
    for (int64_t m = 0; m < nmat; m++)
        if (row_off[m + 1] < row_off[m] || col_off[m + 1] < col_off[m])
            return vnec_fail("vnec_batch_import_ext: offsets not monotone at member %ld", (long)m);
    SpB_Info rc = SpB_Matrix_import_csr_ext(A, ptr, NULL, val);
    if (rc != SpB_SUCCESS) return rc;
    for (int64_t m = 0; m < nmat; m++) {
        int64_t c0 = col_off[m], nc = col_off[m + 1] - col_off[m];
        for (int64_t j = ptr[row_off[m]]; j < ptr[row_off[m + 1]]; j++) {
            if (col[j] < 0 || col[j] >= nc) {
                a->nnz = 0;
                return vnec_fail("vnec_batch_import_ext: member %ld column %d out of range", (long)m, col[j]);
            }
            a->indices[j] = (SpB_Index)(c0 + col[j]);
        }
    }
    return SpB_SUCCESS;
}

/* Device mirror: int32 row offsets and columns, FP64 values.  An FP32
 * matrix's values are widened exactly (FP32 -> FP64 loses nothing): the
 * layout builders are shared by both value types, and FP32 plans narrow
 * their own value arrays back to FP32, again exactly. */
SpB_Info spb_matrix_upload(spb_csr *a)
{
    if (a->dev_ready) return SpB_SUCCESS;
    int64_t nr = (int64_t)a->base.row, nnz = (int64_t)a->nnz;
    if ((int64_t)a->ptr[nr] != nnz) return vnec_fail("matrix row pointers do not end at nnz");
    if (nnz + nr >= INT32_MAX) return vnec_fail("nnz + nrows must stay below 2^31");
    int32_t *hp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nr + 1));
    int32_t *hc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
    if (!hp || !hc) {
        free(hp); free(hc);
        return vnec_fail("device mirror: out of host memory");
    }
    /* Every kernel indexes x by these columns and walks these row pointers:
     * refuse a structure that would make them read out of bounds. */
    int64_t nc = (int64_t)a->base.col;
    if (a->ptr[0] != 0) {
        free(hp);
        free(hc);
        return vnec_fail("matrix row pointers must start at 0");
    }
    for (int64_t r = 0; r <= nr; r++) {
        if (r > 0 && a->ptr[r] < a->ptr[r - 1]) {
            free(hp);
            free(hc);
            return vnec_fail("matrix row pointers decrease at row %lld", (long long)r);
        }
        hp[r] = (int32_t)a->ptr[r];
    }
    for (int64_t j = 0; j < nnz; j++) {
        if (a->indices[j] >= (SpB_Index)nc) {
            free(hp);
            free(hc);
            return vnec_fail("column index %llu at entry %lld is not below ncols = %lld",
                             (unsigned long long)a->indices[j], (long long)j, (long long)nc);
        }
        hc[j] = (int32_t)a->indices[j];
    }
    if (!vnec_runtime_ready()) {
        free(hp);
        free(hc);
        return SpB_FAILURE;
    }
    a->d_ptr = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nr + 1), NULL);
    a->d_col = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)(nnz ? nnz : 1), NULL);
    a->d_val = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(nnz ? nnz : 1), NULL);
    SpB_Info rc = SpB_SUCCESS;
    if (!a->d_ptr || !a->d_col || !a->d_val) rc = SpB_FAILURE;
    hipStream_t s = vnec_stream();
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(a->d_ptr, hp, sizeof(int32_t) * (size_t)(nr + 1),
                                        hipMemcpyHostToDevice, s), "upload ptr");
    if (rc == SpB_SUCCESS && nnz)
        rc = vnec_hip_ok(hipMemcpyAsync(a->d_col, hc, sizeof(int32_t) * (size_t)nnz,
                                        hipMemcpyHostToDevice, s), "upload col");
    if (rc == SpB_SUCCESS && nnz && a->base.element_type == SpB_FP64)
        rc = vnec_hip_ok(hipMemcpyAsync(a->d_val, a->val, sizeof(double) * (size_t)nnz,
                                        hipMemcpyHostToDevice, s), "upload val");
    if (rc == SpB_SUCCESS && nnz && a->base.element_type == SpB_FP32) {
        float *tmp = (float *)vnec_malloc_dev(sizeof(float) * (size_t)nnz, NULL);
        rc = tmp ? SpB_SUCCESS : SpB_FAILURE;
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(hipMemcpyAsync(tmp, a->val, sizeof(float) * (size_t)nnz,
                                            hipMemcpyHostToDevice, s), "upload val");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_widen_f32(tmp, nnz, a->d_val, s), "widen");
        if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "upload sync");
        vnec_free_dev(tmp);
    }
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "upload sync");
    free(hp);
    free(hc);
    if (rc != SpB_SUCCESS) {
        matrix_drop_device(a);
        return SpB_FAILURE;
    }
    a->dev_ready = 1;
    return SpB_SUCCESS;
}

/* ---------------------------------------------------------------- vector */
spb_dense *spb_as_dense(const SpB_Vector v)
{
    spb_dense *d = (spb_dense *)v;
    if (!d || d->magic != SPB_MAGIC_VECTOR) {
        vnec_fail("not a vector made by SpB_Vector_new");
        return NULL;
    }
    return d;
}

/* Host storage of a vector's values: page-locked when a GPU is present, so
 * the uploads of x and downloads of y move by DMA at the link's speed
 * instead of through the runtime's staging copies; else 64-byte aligned
 * pageable memory.  Either way 64-byte aligned host memory. */
static void *vector_host_alloc(size_t bytes, int *pinned)
{
    *pinned = 0;
    if (bytes == 0) bytes = 64;
    if (vnec_runtime_ready()) {
        void *p = NULL;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess && p) {
            *pinned = 1;
            return p;
        }
        (void)hipGetLastError(); /* fall back to pageable memory */
    } else {
        vnec_clear_error(); /* no device: host memory only */
    }
    return aligned64(bytes);
}

SpB_Info SpB_Vector_new(SpB_Vector *v, SpB_Type type, SpB_Index n)
{
    if (!v) return vnec_fail("SpB_Vector_new: NULL handle");
    if (type != SpB_FP64 && type != SpB_FP32)
        return vnec_fail("SpB_Vector_new: the VNEC path takes SpB_FP64 or SpB_FP32 values");
    spb_dense *d = (spb_dense *)calloc(1, sizeof(spb_dense));
    if (!d) return vnec_fail("SpB_Vector_new: out of memory");
    d->base.element_type = type;
    d->base.format_type = SpB_DENSE;
    d->base.n = n;
    d->base.element_size = type == SpB_FP32 ? sizeof(float) : sizeof(double);
    d->values = vector_host_alloc(d->base.element_size * (size_t)n, &d->pinned);
    if (!d->values) {
        free(d);
        return vnec_fail("SpB_Vector_new: out of memory");
    }
    memset(d->values, 0, d->base.element_size * (size_t)n);
    d->magic = SPB_MAGIC_VECTOR;
    d->dev_stale = 1;
    *v = (SpB_Vector)d;
    return SpB_SUCCESS;
}

SpB_Info SpB_Vector_free(SpB_Vector *v)
{
    if (!v || !*v) return vnec_fail("SpB_Vector_free: NULL handle");
    spb_dense *d = (spb_dense *)*v;
    if (d->magic != SPB_MAGIC_VECTOR) return vnec_fail("SpB_Vector_free: not a vector");
    vnec_free_dev(d->d_values);
    if (d->pinned)
        (void)hipHostFree(d->values);
    else
        free(d->values);
    d->magic = 0;
    free(d);
    *v = NULL;
    return SpB_SUCCESS;
}

SpB_Info SpB_Vector_nvals(SpB_Index *nvals, SpB_Vector v)
{
    spb_dense *d = spb_as_dense(v);
    if (!d || !nvals) return SpB_FAILURE;
    *nvals = d->base.n;
    return SpB_SUCCESS;
}

SpB_Info spb_vector_device_alloc(spb_dense *v)
{
    if (v->d_values) return SpB_SUCCESS;
    if (!vnec_runtime_ready()) return SpB_FAILURE;
    v->d_values = vnec_malloc_dev(v->base.element_size * (size_t)v->base.n, NULL);
    if (!v->d_values) return SpB_FAILURE;
    v->dev_stale = 1;
    return SpB_SUCCESS;
}

SpB_Info spb_vector_to_device(spb_dense *v)
{
    if (spb_vector_device_alloc(v) != SpB_SUCCESS) return SpB_FAILURE;
    if (!v->resident || v->dev_stale) {
        if (v->host_stale) return SpB_SUCCESS; /* device already newer */
        SpB_Info rc = vnec_hip_ok(hipMemcpyAsync(v->d_values, v->values,
                                                 v->base.element_size * (size_t)v->base.n,
                                                 hipMemcpyHostToDevice, vnec_stream()),
                                  "vector upload");
        if (rc != SpB_SUCCESS) return rc;
        v->dev_stale = 0;
    }
    return SpB_SUCCESS;
}

SpB_Info spb_vector_from_device(spb_dense *v)
{
    if (!v->host_stale) return SpB_SUCCESS;
    SpB_Info rc = vnec_hip_ok(hipMemcpyAsync(v->values, v->d_values,
                                             v->base.element_size * (size_t)v->base.n,
                                             hipMemcpyDeviceToHost, vnec_stream()),
                              "vector download");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(vnec_stream()), "download sync");
    if (rc == SpB_SUCCESS) v->host_stale = 0;
    return rc;
}

static spb_dense *typed_vector(const SpB_Vector v, SpB_Type type, const char *fn)
{
    spb_dense *d = spb_as_dense(v);
    if (!d) return NULL;
    if (d->base.element_type != type) {
        vnec_fail("%s: the vector holds %s values", fn, d->base.element_type == SpB_FP32 ? "FP32" : "FP64");
        return NULL;
    }
    return d;
}

SpB_Info SpB_Vector_setElement_FP64(SpB_Vector w, double x, SpB_Index i)
{
    spb_dense *d = typed_vector(w, SpB_FP64, "SpB_Vector_setElement_FP64");
    if (!d) return SpB_FAILURE;
    if (i >= d->base.n) return vnec_fail("SpB_Vector_setElement_FP64: index %lu >= %lu", i, d->base.n);
    if (d->host_stale && spb_vector_from_device(d) != SpB_SUCCESS) return SpB_FAILURE;
    ((double *)d->values)[i] = x;
    d->dev_stale = 1;
    return SpB_SUCCESS;
}

SpB_Info SpB_Vector_setElement_FP32(SpB_Vector w, float x, SpB_Index i)
{
    spb_dense *d = typed_vector(w, SpB_FP32, "SpB_Vector_setElement_FP32");
    if (!d) return SpB_FAILURE;
    if (i >= d->base.n) return vnec_fail("SpB_Vector_setElement_FP32: index %lu >= %lu", i, d->base.n);
    if (d->host_stale && spb_vector_from_device(d) != SpB_SUCCESS) return SpB_FAILURE;
    ((float *)d->values)[i] = x;
    d->dev_stale = 1;
    return SpB_SUCCESS;
}

SpB_Info SpB_Vector_extractElement_FP64(double *x, const SpB_Vector v, SpB_Index i)
{
    spb_dense *d = typed_vector(v, SpB_FP64, "SpB_Vector_extractElement_FP64");
    if (!d || !x) return SpB_FAILURE;
    if (i >= d->base.n) return vnec_fail("SpB_Vector_extractElement_FP64: index out of range");
    if (d->host_stale && spb_vector_from_device(d) != SpB_SUCCESS) return SpB_FAILURE;
    *x = ((double *)d->values)[i];
    return SpB_SUCCESS;
}

SpB_Info SpB_Vector_extractElement_FP32(float *x, const SpB_Vector v, SpB_Index i)
{
    spb_dense *d = typed_vector(v, SpB_FP32, "SpB_Vector_extractElement_FP32");
    if (!d || !x) return SpB_FAILURE;
    if (i >= d->base.n) return vnec_fail("SpB_Vector_extractElement_FP32: index out of range");
    if (d->host_stale && spb_vector_from_device(d) != SpB_SUCCESS) return SpB_FAILURE;
    *x = ((float *)d->values)[i];
    return SpB_SUCCESS;
}

/* All n entries of a dense vector as (index, value) tuples
 * (SpB_Vector_Dense.c:162-184): *p_nvals must equal n, else SpB_FAILURE;
 * I[i] = i and X[i] = v[i].  Either array may be NULL to skip it.  The
 * vector's type must match the call's. */
#define SPB_EXTRACT_TUPLES(T, ctype, spbtype)                                                     \
    SpB_Info SpB_Vector_extractTuples_##T(SpB_Index *I, ctype *X, SpB_Index *p_nvals,             \
                                          const SpB_Vector v)                                     \
    {                                                                                             \
        spb_dense *d = typed_vector(v, spbtype, "SpB_Vector_extractTuples_" #T);                   \
        if (!d || !p_nvals) return SpB_FAILURE;                                                   \
        if (*p_nvals != d->base.n)                                                                \
            return vnec_fail("SpB_Vector_extractTuples_" #T ": *p_nvals differs from the size");  \
        if (X && d->host_stale && spb_vector_from_device(d) != SpB_SUCCESS) return SpB_FAILURE;   \
        for (SpB_Index i = 0; i < d->base.n; i++) {                                               \
            if (I) I[i] = i;                                                                      \
            if (X) X[i] = ((const ctype *)d->values)[i];                                          \
        }                                                                                         \
        return SpB_SUCCESS;                                                                       \
    }
SPB_EXTRACT_TUPLES(FP64, double, SpB_FP64)
SPB_EXTRACT_TUPLES(FP32, float, SpB_FP32)
#undef SPB_EXTRACT_TUPLES

SpB_Info SpB_Vector_import_ext(SpB_Vector v, const double *values)
{
    spb_dense *d = typed_vector(v, SpB_FP64, "SpB_Vector_import_ext");
    if (!d || !values) return SpB_FAILURE;
    memcpy(d->values, values, sizeof(double) * (size_t)d->base.n);
    d->host_stale = 0;
    d->dev_stale = 1;
    return SpB_SUCCESS;
}

SpB_Info SpB_Vector_export_ext(const SpB_Vector v, double *values)
{
    spb_dense *d = typed_vector(v, SpB_FP64, "SpB_Vector_export_ext");
    if (!d || !values) return SpB_FAILURE;
    if (d->host_stale && spb_vector_from_device(d) != SpB_SUCCESS) return SpB_FAILURE;
    memcpy(values, d->values, sizeof(double) * (size_t)d->base.n);
    return SpB_SUCCESS;
}

SpB_Info SpB_Vector_import_fp32_ext(SpB_Vector v, const float *values)
{
    spb_dense *d = typed_vector(v, SpB_FP32, "SpB_Vector_import_fp32_ext");
    if (!d || !values) return SpB_FAILURE;
    memcpy(d->values, values, sizeof(float) * (size_t)d->base.n);
    d->host_stale = 0;
    d->dev_stale = 1;
    return SpB_SUCCESS;
}

SpB_Info SpB_Vector_export_fp32_ext(const SpB_Vector v, float *values)
{
    spb_dense *d = typed_vector(v, SpB_FP32, "SpB_Vector_export_fp32_ext");
    if (!d || !values) return SpB_FAILURE;
    if (d->host_stale && spb_vector_from_device(d) != SpB_SUCCESS) return SpB_FAILURE;
    memcpy(values, d->values, sizeof(float) * (size_t)d->base.n);
    return SpB_SUCCESS;
}

SpB_Info SpB_Vector_device_resident_ext(SpB_Vector v, bool on)
{
    spb_dense *d = spb_as_dense(v);
    if (!d) return SpB_FAILURE;
    if (!on && d->host_stale && spb_vector_from_device(d) != SpB_SUCCESS) return SpB_FAILURE;
    d->resident = on ? 1 : 0;
    if (on) return spb_vector_to_device(d);
    return SpB_SUCCESS;
}

SpB_Info SpB_Vector_sync_host_ext(SpB_Vector v)
{
    spb_dense *d = spb_as_dense(v);
    if (!d) return SpB_FAILURE;
    return spb_vector_from_device(d);
}

void *SpB_Vector_device_ptr_ext(SpB_Vector v)
{
    spb_dense *d = spb_as_dense(v);
    if (!d) return NULL;
    if (spb_vector_device_alloc(d) != SpB_SUCCESS) return NULL;
    return d->d_values;
}

/* ------------------------------------------------- device memory helpers */
void *vnec_dev_alloc_ext(size_t bytes)
{
    if (!vnec_runtime_ready()) return NULL;
    return vnec_malloc_dev(bytes, NULL);
}

SpB_Info vnec_dev_free_ext(void *p)
{
    if (p) return vnec_hip_ok(hipFree(p), "hipFree");
    return SpB_SUCCESS;
}

SpB_Info vnec_memcpy_h2d_ext(void *dst, const void *src, size_t bytes)
{
    if (!vnec_runtime_ready()) return SpB_FAILURE;
    SpB_Info rc = vnec_hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, vnec_stream()),
                              "h2d");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(vnec_stream()), "h2d sync");
    return rc;
}

SpB_Info vnec_memcpy_d2h_ext(void *dst, const void *src, size_t bytes)
{
    if (!vnec_runtime_ready()) return SpB_FAILURE;
    SpB_Info rc = vnec_hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, vnec_stream()),
                              "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(vnec_stream()), "d2h sync");
    return rc;
}

SpB_Info vnec_sync_ext(void)
{
    if (!vnec_runtime_ready()) return SpB_FAILURE;
    return vnec_hip_ok(hipStreamSynchronize(vnec_stream()), "sync");
}
