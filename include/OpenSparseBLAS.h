/*
 * OpenSparseBLAS.h — the drop-in subset of the OpenSparseBLAS C API that the
 * VNEC SpMV path uses (FP64, and its FP32 twin), implemented by
 * libvnec_mi355x.so on one MI355X.
 *
 * Every declaration keeps the reference's name, signature, enum values and
 * public struct layout; each cites the line of the reference header
 * (include/OpenSparseBLAS.h in the VNEC repository) that it replaces.  Only
 * the path's subset is declared: the object life cycle, FP64 / FP32 build,
 * set/extract, SpB_mxv and the VNEC plan API.
 *
 * Behavioural notes (where this library differs on purpose):
 *   - Errors return SpB_FAILURE (or NULL for builders) and never exit();
 *     vnec_last_error() (vnec_ext.h) gives the reason.
 *   - SpB_Matrix_build_FP64 fills every row pointer, including trailing empty
 *     rows (the reference leaves them at 0).
 *   - Plans own device memory: release them with the *_free_ext calls of
 *     vnec_ext.h.  Freeing the public fields field by field (as the
 *     reference harness does) still works for the host arrays.
 */
#pragma once
#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/OpenSparseBLAS.h:32-37 */
typedef enum {
    SpB_SUCCESS = 0,
    SpB_FAILURE = 1,
    SpB_NO_VALUE = -1
} SpB_Info;

/* include/OpenSparseBLAS.h:42-56 */
typedef enum {
    SpB_NULL_TYPE,
    SpB_BOOL,
    SpB_INT8,
    SpB_INT16,
    SpB_INT32,
    SpB_INT64,
    SpB_UINT8,
    SpB_UINT16,
    SpB_UINT32,
    SpB_UINT64,
    SpB_FP32,
    SpB_FP64,
    SpB_Type_N
} SpB_Type;

/* include/OpenSparseBLAS.h:78-139: every enumerator, with the reference's
 * values (0, 1, 2, ... in this order), so code naming any of them compiles
 * against this header.  SpB_Matrix_build_* accepts only SpB_BINARYOP_NULL
 * (duplicates kept) and returns SpB_FAILURE for any other operator, as the
 * reference's build does (src/core/src/SpB_Matrix_CSC_or_CSR.c:197,252). */
typedef enum {
    SpB_BINARYOP_NULL,
    SpB_BINARYOP_BOOL_ST,
    SpB_LOR,
    SpB_LAND,
    SpB_LXOR,
    SpB_LXNOR,
    SpB_BINARYOP_BOOL_ED,
    SpB_BINARYOP_INT8_ST,
    SpB_PLUS_INT8,
    SpB_MINUS_INT8,
    SpB_DIV_INT8,
    SpB_BINARYOP_INT8_ED,
    SpB_BINARYOP_INT16_ST,
    SpB_PLUS_INT16,
    SpB_MINUS_INT16,
    SpB_DIV_INT16,
    SpB_BINARYOP_INT16_ED,
    SpB_BINARYOP_INT32_ST,
    SpB_PLUS_INT32,
    SpB_MINUS_INT32,
    SpB_DIV_INT32,
    SpB_BINARYOP_INT32_ED,
    SpB_BINARYOP_INT64_ST,
    SpB_PLUS_INT64,
    SpB_MINUS_INT64,
    SpB_DIV_INT64,
    SpB_BINARYOP_INT64_ED,
    SpB_BINARYOP_UINT8_ST,
    SpB_PLUS_UINT8,
    SpB_MINUS_UINT8,
    SpB_DIV_UINT8,
    SpB_BINARYOP_UINT8_ED,
    SpB_BINARYOP_UINT16_ST,
    SpB_PLUS_UINT16,
    SpB_MINUS_UINT16,
    SpB_DIV_UINT16,
    SpB_BINARYOP_UINT16_ED,
    SpB_BINARYOP_UINT32_ST,
    SpB_PLUS_UINT32,
    SpB_MINUS_UINT32,
    SpB_DIV_UINT32,
    SpB_BINARYOP_UINT32_ED,
    SpB_BINARYOP_UINT64_ST,
    SpB_PLUS_UINT64,
    SpB_MINUS_UINT64,
    SpB_DIV_UINT64,
    SpB_BINARYOP_UINT64_ED,
    SpB_BINARYOP_FP32_ST,
    SpB_PLUS_FP32,
    SpB_TIMES_FP32,
    SpB_DIV_FP32,
    SpB_MAX_FP32,
    SpB_MINUS_FP32,
    SpB_BINARYOP_FP32_ED,
    SpB_BINARYOP_FP64_ST,
    SpB_PLUS_FP64,
    SpB_MINUS_FP64,
    SpB_DIV_FP64,
    SpB_BINARYOP_FP64_ED
} SpB_BinaryOp;

/* include/OpenSparseBLAS.h:252 */
typedef unsigned long int SpB_Index;

/* include/OpenSparseBLAS.h:268, 273 */
void SpB_init();
void SpB_finalize();

/* include/OpenSparseBLAS.h:318-323, 946-949 */
struct SpB_Vector_opaque;
typedef struct SpB_Vector_opaque *SpB_Vector;
struct SpB_Matrix_opaque;
typedef struct SpB_Matrix_opaque *SpB_Matrix;

/* include/OpenSparseBLAS.h:339, 347, 361 */
SpB_Info SpB_Vector_new(SpB_Vector *v, SpB_Type type, SpB_Index n);
SpB_Info SpB_Vector_free(SpB_Vector *v);
SpB_Info SpB_Vector_nvals(SpB_Index *nvals, SpB_Vector v);

/* include/OpenSparseBLAS.h:538-543, 549-554 */
extern SpB_Info SpB_Vector_setElement_FP32(SpB_Vector w, float x, SpB_Index i);
extern SpB_Info SpB_Vector_setElement_FP64(SpB_Vector w, double x, SpB_Index i);
/* include/OpenSparseBLAS.h:730-735, 741-746 */
extern SpB_Info SpB_Vector_extractElement_FP32(float *x, const SpB_Vector v, SpB_Index i);
extern SpB_Info SpB_Vector_extractElement_FP64(double *x, const SpB_Vector v, SpB_Index i);
/* All n entries as (i, v[i]) (SpB_Vector.c:423-457, SpB_Vector_Dense.c:162-184):
 * *p_nvals must equal n; I or X may be NULL. */
extern SpB_Info SpB_Vector_extractTuples_FP32(SpB_Index *I, float *X, SpB_Index *p_nvals,
                                              const SpB_Vector v);
extern SpB_Info SpB_Vector_extractTuples_FP64(SpB_Index *I, double *X, SpB_Index *p_nvals,
                                              const SpB_Vector v);

/* include/OpenSparseBLAS.h:971, 981, 992, 1000 */
SpB_Info SpB_Matrix_nrows(SpB_Index *n, SpB_Matrix m);
SpB_Info SpB_Matrix_ncols(SpB_Index *n, SpB_Matrix m);
SpB_Info SpB_Matrix_new(SpB_Matrix *m, SpB_Type type, SpB_Index nrows, SpB_Index ncols);
SpB_Info SpB_Matrix_free(SpB_Matrix *A);

/* include/OpenSparseBLAS.h:1152-1160 */
extern SpB_Info SpB_Matrix_build_FP32(SpB_Matrix C, const SpB_Index *I, const SpB_Index *J,
                                      const float *X, SpB_Index nvals, const SpB_BinaryOp dup);
/* include/OpenSparseBLAS.h:1166-1174 */
extern SpB_Info SpB_Matrix_build_FP64(SpB_Matrix C, const SpB_Index *I, const SpB_Index *J,
                                      const double *X, SpB_Index nvals, const SpB_BinaryOp dup);

/* include/OpenSparseBLAS.h:1951 — declared but never defined in the
 * reference; here y <- A x (y overwritten).  VNEC is a plan returned by a
 * Build_VNEC_* call or by vnec_plan_build_ext, or NULL to build (with the
 * learned selector) and cache one on A. */
SpB_Info SpB_mxv(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, void *VNEC);

/* include/OpenSparseBLAS.h:1954-1958 */
typedef enum {
    SpB_VNEC_D = 0,
    SpB_VNEC_S = 1,
} SpB_VNEC_type;

/* include/OpenSparseBLAS.h:1960-1964 */
typedef struct coord {
    int x;
    int y;
} coord;

/* include/OpenSparseBLAS.h:1983 */
SpB_VNEC_type IRD_VNEC(const SpB_Matrix A, SpB_Type type);

/* include/OpenSparseBLAS.h:1966-1982 (FP32 values, x, y and sums) */
typedef struct VNEC_D_FP32 {
    coord *thread_coord_start;
    coord *thread_coord_end;
    int *ecr_indices;
    int *NEC_NUM;
    int **use_x_indices;
    float **ecr_xx_val;
    int *v_row_ptr;
    int *col_start;
    float *val_align;
} VNEC_D_FP32;
void SpMV_VNEC_D_FP32(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_D_FP32 *mat_thd);
struct VNEC_D_FP32 *Build_VNEC_D_FP32(const SpB_Matrix A);

/* include/OpenSparseBLAS.h:1985-2012 */
struct csr_f32 {
    float *nnz;
    int *col, *row_begin, *row_end;
    int *task_start;
    int *task_end;
};
typedef struct VNEC_S_FP32 {
    int *NEC_NUM;
    int **use_x_indices;
    float **ecr_xx_val;
    int *spvv16_len;
    int **tasks;
    struct csr_f32 reorder_mat;
} VNEC_S_FP32;
void SpMV_VNEC_S_FP32(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_S_FP32 *mat_thd);
struct VNEC_S_FP32 *Build_VNEC_S_FP32(const SpB_Matrix A);

/* include/OpenSparseBLAS.h:2014-2023 */
typedef struct VNEC_L {
    int num_merge_items;
    int items_per_thread;
    int *diagonal_start;
    int *diagonal_end;
    coord *thread_coord_start;
    coord *thread_coord_end;
    int *nz_indices;
} VNEC_L;
struct VNEC_L *Build_VNEC_L(const SpB_Matrix A);
/* include/OpenSparseBLAS.h:2024 */
void SpMV_VNEC_L_FP32(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_L *mat_thd);

/* include/OpenSparseBLAS.h:2027-2043 */
typedef struct VNEC_D_FP64 {
    coord *thread_coord_start;
    coord *thread_coord_end;
    int *ecr_indices;
    int *NEC_NUM;
    int **use_x_indices;
    double **ecr_xx_val;
    int *v_row_ptr;
    int *col_start;
    double *val_align;
} VNEC_D_FP64;
void SpMV_VNEC_D_FP64(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_D_FP64 *mat_thd);
struct VNEC_D_FP64 *Build_VNEC_D_FP64(const SpB_Matrix A);

/* include/OpenSparseBLAS.h:2044-2071 */
struct csr_f64 {
    double *nnz;
    int *col, *row_begin, *row_end;
    int *task_start;
    int *task_end;
};
typedef struct VNEC_S_FP64 {
    int *NEC_NUM;
    int **use_x_indices;
    double **ecr_xx_val;
    int *spvv8_len;
    int **tasks;
    struct csr_f64 reorder_mat;
} VNEC_S_FP64;
void SpMV_VNEC_S_FP64(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_S_FP64 *mat_thd);
struct VNEC_S_FP64 *Build_VNEC_S_FP64(const SpB_Matrix A);

/* include/OpenSparseBLAS.h:2073 */
void SpMV_VNEC_L_FP64(SpB_Vector y, const SpB_Matrix A, const SpB_Vector x, VNEC_L *mat_thd);

/* include/OpenSparseBLAS.h:2532-2581 */
typedef enum {
    SpB_VECTOR_FORMAT,
    SpB_MATRIX_FORMAT
} SpB_Option_Field_ext;

typedef enum {
    SpB_DENSE = 0,
    SpB_SPARSE = 1,
    SpB_VECTOR_NO_FORMAT = -1
} SpB_Vector_Format_Value_ext;

typedef enum {
    SpB_CSC = 0,
    SpB_CSR = 1,
    SpB_MATRIX_NO_FORMAT = -1
} SpB_Matrix_Format_Value_ext;

typedef union {
    SpB_Vector_Format_Value_ext vector_format;
    SpB_Matrix_Format_Value_ext matrix_format;
} SpB_Object_Format_ext;

SpB_Info SpB_Global_Option_set_ext(SpB_Option_Field_ext field, SpB_Object_Format_ext value);

#ifdef __cplusplus
}

/* C++ overloads by value type (OpenSparseBLAS.h:2651-2701 of the reference,
 * there defined in SpB_Vector.cpp), for the value types this library's
 * vectors hold; header-only, so the library itself stays C. */
inline SpB_Info SpB_Vector_setElement(SpB_Vector v, double x, SpB_Index i)
{
    return SpB_Vector_setElement_FP64(v, x, i);
}
inline SpB_Info SpB_Vector_setElement(SpB_Vector v, float x, SpB_Index i)
{
    return SpB_Vector_setElement_FP32(v, x, i);
}
inline SpB_Info SpB_Vector_extractElement(double *x, const SpB_Vector v, SpB_Index i)
{
    return SpB_Vector_extractElement_FP64(x, v, i);
}
inline SpB_Info SpB_Vector_extractElement(float *x, const SpB_Vector v, SpB_Index i)
{
    return SpB_Vector_extractElement_FP32(x, v, i);
}
inline SpB_Info SpB_Vector_extractTuples(SpB_Index *I, double *X, SpB_Index *p_nvals, const SpB_Vector v)
{
    return SpB_Vector_extractTuples_FP64(I, X, p_nvals, v);
}
inline SpB_Info SpB_Vector_extractTuples(SpB_Index *I, float *X, SpB_Index *p_nvals, const SpB_Vector v)
{
    return SpB_Vector_extractTuples_FP32(I, X, p_nvals, v);
}
#endif
