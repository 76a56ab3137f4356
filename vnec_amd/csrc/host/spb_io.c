/*
 * spb_io.c — graph file loaders of the reference harness, restated in C:
 * MatrixMarket coordinate files (.mtx) and serialized CSR graphs (.sg).
 *
 * .mtx (utilities/reader.hpp:27-153): header "%%MatrixMarket matrix
 * coordinate <field> <symmetry>" (field pattern / real / double / integer;
 * symmetry general / skew-symmetric / symmetric), '%' comment lines, then
 * "rows cols entries" (the matrix must be square) and one "u v [w]" line
 * per entry, 1-based.  A symmetric file also yields each reversed edge.
 * The edge list is then squished as utilities/builder.hpp:165-187 does:
 * sorted by (u, v), exact duplicates dropped, self-loops removed, and laid
 * out as CSR (columns ascending within each row).
 *
 * .sg (utilities/reader.hpp:189-240, writer.hpp:12-49): bool directed,
 * uint64 edges, uint64 nodes, uint64 offsets[nodes + 1], uint64
 * destinations[edges] (then the incoming arrays when directed, not needed
 * here); the stored CSR is taken as is.
 *
 * Every error returns SpB_FAILURE with a reason (the reference exits).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spb_internal.h"

static int ends_with(const char *s, const char *suf)
{
    size_t a = strlen(s), b = strlen(suf);
    return a >= b && strcmp(s + a - b, suf) == 0;
}

static int cmp_edge(const void *p, const void *q)
{
    const uint64_t a = *(const uint64_t *)p, b = *(const uint64_t *)q;
    return (a > b) - (a < b);
}

/* Edges as (u << 32 | v) keys: sorting the keys sorts by (u, v). */
static SpB_Info squish_to_csr(uint64_t *e, int64_t ne, int64_t n, int64_t **ptr_out, int32_t **col_out)
{
    qsort(e, (size_t)ne, sizeof(uint64_t), cmp_edge);
    int64_t m = 0;
    for (int64_t i = 0; i < ne; i++) {
        if (i > 0 && e[i] == e[i - 1]) continue;       /* duplicate edge */
        if ((e[i] >> 32) == (e[i] & 0xFFFFFFFFull)) continue; /* self-loop */
        e[m++] = e[i];
    }
    int64_t *ptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m ? m : 1));
    if (!ptr || !col) {
        free(ptr);
        free(col);
        return vnec_fail("graph loader: out of memory");
    }
    for (int64_t i = 0; i < m; i++) {
        ptr[(e[i] >> 32) + 1]++;
        col[i] = (int32_t)(e[i] & 0xFFFFFFFFull);
    }
    for (int64_t r = 0; r < n; r++) ptr[r + 1] += ptr[r];
    *ptr_out = ptr;
    *col_out = col;
    return SpB_SUCCESS;
}

static SpB_Info read_mtx(FILE *f, const char *path, int64_t *n_out, int64_t **ptr, int32_t **col)
{
    char start[64], object[64], format[64], field[64], symmetry[64];
    if (fscanf(f, "%63s %63s %63s %63s %63s", start, object, format, field, symmetry) != 5)
        return vnec_fail("%s: missing MatrixMarket header", path);
    if (strcmp(start, "%%MatrixMarket") != 0 && strcmp(start, "%MatrixMarket") != 0)
        return vnec_fail("%s: .mtx file did not start with %%%%MatrixMarket", path);
    if (strcmp(object, "matrix") != 0 || strcmp(format, "coordinate") != 0)
        return vnec_fail("%s: only matrix coordinate .mtx files are supported", path);
    if (strcmp(field, "pattern") && strcmp(field, "real") && strcmp(field, "double") &&
        strcmp(field, "integer"))
        return vnec_fail("%s: unsupported .mtx field '%s'", path, field);
    int symmetric;
    if (strcmp(symmetry, "symmetric") == 0)
        symmetric = 1;
    else if (strcmp(symmetry, "general") == 0 || strcmp(symmetry, "skew-symmetric") == 0)
        symmetric = 0;
    else
        return vnec_fail("%s: unsupported .mtx symmetry '%s'", path, symmetry);
    char line[1024];
    if (!fgets(line, sizeof line, f)) return vnec_fail("%s: truncated header", path); /* rest of line 1 */
    long long m = -1, nn = -1, nz = -1;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '%') continue;
        if (sscanf(line, "%lld %lld %lld", &m, &nn, &nz) == 3) break;
        int blank = 1;
        for (char *c = line; *c; c++)
            if (*c != ' ' && *c != '\t' && *c != '\n' && *c != '\r') blank = 0;
        if (!blank) return vnec_fail("%s: bad size line", path);
    }
    if (m < 0 || nn < 0 || nz < 0) return vnec_fail("%s: missing size line", path);
    if (m != nn) return vnec_fail("%s: matrix must be square for .mtx (%lld x %lld)", path, m, nn);
    if (m >= INT32_MAX) return vnec_fail("%s: too many rows", path);
    int64_t cap = (symmetric ? 2 : 1) * (int64_t)nz, ne = 0;
    uint64_t *e = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(cap ? cap : 1));
    if (!e) return vnec_fail("%s: out of memory", path);
    while (fgets(line, sizeof line, f)) {
        long long u, v;
        if (line[0] == '%') continue;
        if (sscanf(line, "%lld %lld", &u, &v) != 2) continue; /* blank line */
        if (u < 1 || v < 1 || u > m || v > m) {
            free(e);
            return vnec_fail("%s: node id out of range (%lld, %lld)", path, u, v);
        }
        if (ne + (symmetric ? 2 : 1) > cap) {
            free(e);
            return vnec_fail("%s: more entries than the size line declares", path);
        }
        e[ne++] = ((uint64_t)(u - 1) << 32) | (uint64_t)(v - 1);
        if (symmetric) e[ne++] = ((uint64_t)(v - 1) << 32) | (uint64_t)(u - 1);
    }
    SpB_Info rc = squish_to_csr(e, ne, m, ptr, col);
    free(e);
    if (rc == SpB_SUCCESS) *n_out = m;
    return rc;
}

static SpB_Info read_sg(FILE *f, const char *path, int64_t *n_out, int64_t **ptr_out, int32_t **col_out)
{
    unsigned char directed;
    uint64_t ne, n;
    if (fread(&directed, 1, 1, f) != 1 || fread(&ne, 8, 1, f) != 1 || fread(&n, 8, 1, f) != 1)
        return vnec_fail("%s: truncated .sg header", path);
    if (n >= INT32_MAX || ne >= INT32_MAX) return vnec_fail("%s: graph too large for int32 indices", path);
    uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n + 1));
    uint64_t *dst = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(ne ? ne : 1));
    int64_t *ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ne ? ne : 1));
    SpB_Info rc = (off && dst && ptr && col) ? SpB_SUCCESS : vnec_fail("%s: out of memory", path);
    if (rc == SpB_SUCCESS && fread(off, 8, (size_t)(n + 1), f) != (size_t)(n + 1))
        rc = vnec_fail("%s: truncated .sg offsets", path);
    if (rc == SpB_SUCCESS && ne && fread(dst, 8, (size_t)ne, f) != (size_t)ne)
        rc = vnec_fail("%s: truncated .sg destinations", path);
    if (rc == SpB_SUCCESS && (off[0] != 0 || off[n] != ne)) rc = vnec_fail("%s: inconsistent .sg offsets", path);
    for (uint64_t i = 0; rc == SpB_SUCCESS && i <= n; i++) {
        if (i > 0 && off[i] < off[i - 1]) rc = vnec_fail("%s: .sg offsets decrease", path);
        ptr[i] = (int64_t)off[i];
    }
    for (uint64_t j = 0; rc == SpB_SUCCESS && j < ne; j++) {
        if (dst[j] >= n) rc = vnec_fail("%s: .sg destination out of range", path);
        col[j] = (int32_t)dst[j];
    }
    (void)directed; /* the incoming arrays that follow are not needed */
    free(off);
    free(dst);
    if (rc != SpB_SUCCESS) {
        free(ptr);
        free(col);
        return rc;
    }
    *n_out = (int64_t)n;
    *ptr_out = ptr;
    *col_out = col;
    return SpB_SUCCESS;
}

SpB_Info vnec_read_graph_ext(const char *path, int64_t *nrows, int64_t **ptr, int32_t **col)
{
    if (!path || !nrows || !ptr || !col) return vnec_fail("vnec_read_graph_ext: bad arguments");
    *ptr = NULL;
    *col = NULL;
    int sg = ends_with(path, ".sg");
    if (!sg && !ends_with(path, ".mtx")) return vnec_fail("%s: unrecognized suffix (.mtx or .sg)", path);
    FILE *f = fopen(path, sg ? "rb" : "r");
    if (!f) return vnec_fail("couldn't open file %s", path);
    SpB_Info rc = sg ? read_sg(f, path, nrows, ptr, col) : read_mtx(f, path, nrows, ptr, col);
    fclose(f);
    return rc;
}

void vnec_host_free_ext(void *p) { free(p); }

SpB_Info SpB_Matrix_read_ext(SpB_Matrix *A, const char *path, SpB_Type type, int values)
{
    if (!A) return vnec_fail("SpB_Matrix_read_ext: NULL handle");
    *A = NULL;
    if (values != VNEC_VALUES_ROW && values != VNEC_VALUES_ONE)
        return vnec_fail("SpB_Matrix_read_ext: values must be VNEC_VALUES_ROW or VNEC_VALUES_ONE");
    if (type != SpB_FP64 && type != SpB_FP32)
        return vnec_fail("SpB_Matrix_read_ext: type must be SpB_FP64 or SpB_FP32");
    int64_t n = 0, *ptr = NULL;
    int32_t *col = NULL;
    if (vnec_read_graph_ext(path, &n, &ptr, &col) != SpB_SUCCESS) return SpB_FAILURE;
    int64_t nnz = ptr[n];
    size_t esz = type == SpB_FP32 ? sizeof(float) : sizeof(double);
    void *val = malloc(esz * (size_t)(nnz ? nnz : 1));
    SpB_Info rc = val ? SpB_SUCCESS : vnec_fail("SpB_Matrix_read_ext: out of memory");
    for (int64_t r = 0; rc == SpB_SUCCESS && r < n; r++)
        for (int64_t j = ptr[r]; j < ptr[r + 1]; j++) {
            double v = values == VNEC_VALUES_ROW ? (double)r : 1.0;
            if (type == SpB_FP32) ((float *)val)[j] = (float)v;
            else ((double *)val)[j] = v;
        }
    SpB_Object_Format_ext fmt;
    fmt.matrix_format = SpB_CSR;
    if (rc == SpB_SUCCESS) rc = SpB_Global_Option_set_ext(SpB_MATRIX_FORMAT, fmt);
    if (rc == SpB_SUCCESS) rc = SpB_Matrix_new(A, type, (SpB_Index)n, (SpB_Index)n);
    if (rc == SpB_SUCCESS)
        rc = type == SpB_FP32 ? SpB_Matrix_import_csr_fp32_ext(*A, ptr, col, (const float *)val)
                              : SpB_Matrix_import_csr_ext(*A, ptr, col, (const double *)val);
    if (rc != SpB_SUCCESS && *A) SpB_Matrix_free(A);
    free(ptr);
    free(col);
    free(val);
    return rc;
}
