/*
 * ref_driver.c — a ctypes-friendly driver around the reference VNEC library
 * compiled from its own sources (see oracle/Makefile, target `ref`).
 *
 * TEST INFRASTRUCTURE ONLY: used to write golden fixtures and, where the
 * host CPU has AVX-512, as the timed CPU baseline.  It builds the
 * reference's CSR object directly from a caller CSR (bypassing
 * SpB_Matrix_build_FP64, whose pointer fill leaves trailing empty rows at
 * 0 — SpB_Matrix_CSC_or_CSR.c:239-249) and copies the plan structures the
 * reference builders return into caller buffers.
 *
 * The value array is padded by 16 doubles because Build_VNEC_D_FP64 reads
 * 8 entries past nnz (spmv_vnec_D_for_mxv.c:339-344).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "SpB_mxv.h"

int ref_num_threads(void) { return NUM_THREADS; }
int ref_lanes_f64(void) { return nLanes_f64; }

static int g_init = 0;

static void ensure_init(void)
{
    if (g_init) return;
    SpB_init();
    SpB_Object_Format_ext f;
    f.matrix_format = SpB_CSR;
    SpB_Global_Option_set_ext(SpB_MATRIX_FORMAT, f);
    g_init = 1;
}

void *ref_matrix(int64_t nrows, int64_t ncols, const int64_t *ptr, const int32_t *col,
                 const double *val)
{
    ensure_init();
    SpB_Matrix A = NULL;
    if (SpB_Matrix_new(&A, SpB_FP64, (SpB_Index)nrows, (SpB_Index)ncols) != SpB_SUCCESS) return NULL;
    SpB_Matrix_CSC_or_CSR *a = (SpB_Matrix_CSC_or_CSR *)A;
    int64_t nnz = ptr[nrows];
    free(a->ptr);
    a->ptr = (SpB_Index *)malloc(sizeof(SpB_Index) * (size_t)(nrows + 1));
    a->indices = (SpB_Index *)malloc(sizeof(SpB_Index) * (size_t)(nnz + 16));
    a->val = calloc((size_t)(nnz + 16), sizeof(double));
    for (int64_t i = 0; i <= nrows; i++) a->ptr[i] = (SpB_Index)ptr[i];
    for (int64_t j = 0; j < nnz; j++) a->indices[j] = (SpB_Index)col[j];
    memcpy(a->val, val, sizeof(double) * (size_t)nnz);
    a->ptr_len = (SpB_Index)(nrows + 1);
    a->nnz = (SpB_Index)nnz;
    a->capacity = (SpB_Index)nnz;
    return A;
}

void ref_matrix_free(void *h)
{
    SpB_Matrix A = (SpB_Matrix)h;
    SpB_Matrix_free(&A);
}

static SpB_Vector make_vec(const double *v, int64_t n)
{
    SpB_Vector x = NULL;
    SpB_Vector_new(&x, SpB_FP64, (SpB_Index)n);
    if (v) memcpy(((SpB_Vector_Dense *)x)->values, v, sizeof(double) * (size_t)n);
    return x;
}

int ref_ird(void *h) { return (int)IRD_VNEC((SpB_Matrix)h, SpB_FP64); }

/* ------------------------------------------------------------------ D */
void *ref_build_d(void *h) { return Build_VNEC_D_FP64((SpB_Matrix)h); }

int64_t ref_d_groups(void *p, void *h)
{
    return ((VNEC_D_FP64 *)p)->v_row_ptr[((SpB_Matrix)h)->row];
}

/* coords: 4 * P ints (start.x, start.y, end.x, end.y per partition);
 * win: concatenated use_x_indices trimmed to NEC_NUM. */
void ref_d_export(void *p, void *h, int32_t *coords, int32_t *ecr, int32_t *nec, int32_t *win,
                  int32_t *v_row_ptr, int32_t *col_start, double *val_align)
{
    VNEC_D_FP64 *d = (VNEC_D_FP64 *)p;
    SpB_Matrix_CSC_or_CSR *a = (SpB_Matrix_CSC_or_CSR *)h;
    int64_t nrows = (int64_t)a->matrix_base.row, nnz = (int64_t)a->nnz;
    int64_t off = 0;
    for (int t = 0; t < NUM_THREADS; t++) {
        coords[4 * t + 0] = d->thread_coord_start[t].x;
        coords[4 * t + 1] = d->thread_coord_start[t].y;
        coords[4 * t + 2] = d->thread_coord_end[t].x;
        coords[4 * t + 3] = d->thread_coord_end[t].y;
        nec[t] = d->NEC_NUM[t];
        memcpy(win + off, d->use_x_indices[t], sizeof(int32_t) * (size_t)d->NEC_NUM[t]);
        off += d->NEC_NUM[t];
    }
    memcpy(ecr, d->ecr_indices, sizeof(int32_t) * (size_t)nnz);
    memcpy(v_row_ptr, d->v_row_ptr, sizeof(int32_t) * (size_t)(nrows + 1));
    int64_t G = d->v_row_ptr[nrows];
    memcpy(col_start, d->col_start, sizeof(int32_t) * (size_t)G);
    memcpy(val_align, d->val_align, sizeof(double) * (size_t)(G * nLanes_f64));
}

void ref_spmv_d(void *p, void *h, const double *x, double *y, int reps)
{
    SpB_Matrix A = (SpB_Matrix)h;
    SpB_Vector xv = make_vec(x, (int64_t)A->col), yv = make_vec(NULL, (int64_t)A->row);
    for (int r = 0; r < reps; r++) SpMV_VNEC_D_FP64(yv, A, xv, (VNEC_D_FP64 *)p);
    memcpy(y, ((SpB_Vector_Dense *)yv)->values, sizeof(double) * A->row);
    SpB_Vector_free(&xv);
    SpB_Vector_free(&yv);
}

void ref_free_d(void *p)
{
    VNEC_D_FP64 *d = (VNEC_D_FP64 *)p;
    free(d->thread_coord_start);
    free(d->thread_coord_end);
    free(d->ecr_indices);
    free(d->NEC_NUM);
    for (int t = 0; t < NUM_THREADS; t++) {
        free(d->use_x_indices[t]);
        _mm_free(d->ecr_xx_val[t]);
    }
    free(d->use_x_indices);
    free(d->ecr_xx_val);
    free(d->col_start);
    free(d->v_row_ptr);
    _mm_free(d->val_align);
    free(d);
}

/* ------------------------------------------------------------------ S */
void *ref_build_s(void *h) { return Build_VNEC_S_FP64((SpB_Matrix)h); }

/* tasks: nrows ints (concatenated per partition); task_se: 2 * P ints. */
void ref_s_export(void *p, void *h, int32_t *nec, int32_t *win, int32_t *spvv_len, int32_t *tasks,
                  int32_t *row_begin, int32_t *row_end, int32_t *s_col, double *s_val,
                  int32_t *task_se)
{
    VNEC_S_FP64 *s = (VNEC_S_FP64 *)p;
    SpB_Matrix_CSC_or_CSR *a = (SpB_Matrix_CSC_or_CSR *)h;
    int64_t nrows = (int64_t)a->matrix_base.row, nnz = (int64_t)a->nnz;
    int64_t off = 0, toff = 0;
    for (int t = 0; t < NUM_THREADS; t++) {
        nec[t] = s->NEC_NUM[t];
        memcpy(win + off, s->use_x_indices[t], sizeof(int32_t) * (size_t)s->NEC_NUM[t]);
        off += s->NEC_NUM[t];
        spvv_len[t] = s->spvv8_len[t];
        int64_t nr = s->reorder_mat.task_end[t] - s->reorder_mat.task_start[t];
        memcpy(tasks + toff, s->tasks[t], sizeof(int32_t) * (size_t)nr);
        toff += nr;
        task_se[2 * t + 0] = s->reorder_mat.task_start[t];
        task_se[2 * t + 1] = s->reorder_mat.task_end[t];
    }
    memcpy(row_begin, s->reorder_mat.row_begin, sizeof(int32_t) * (size_t)nrows);
    memcpy(row_end, s->reorder_mat.row_end, sizeof(int32_t) * (size_t)nrows);
    memcpy(s_col, s->reorder_mat.col, sizeof(int32_t) * (size_t)nnz);
    memcpy(s_val, s->reorder_mat.nnz, sizeof(double) * (size_t)nnz);
}

void ref_spmv_s(void *p, void *h, const double *x, double *y, int reps)
{
    SpB_Matrix A = (SpB_Matrix)h;
    SpB_Vector xv = make_vec(x, (int64_t)A->col), yv = make_vec(NULL, (int64_t)A->row);
    for (int r = 0; r < reps; r++) SpMV_VNEC_S_FP64(yv, A, xv, (VNEC_S_FP64 *)p);
    memcpy(y, ((SpB_Vector_Dense *)yv)->values, sizeof(double) * A->row);
    SpB_Vector_free(&xv);
    SpB_Vector_free(&yv);
}

void ref_free_s(void *p)
{
    VNEC_S_FP64 *s = (VNEC_S_FP64 *)p;
    free(s->NEC_NUM);
    free(s->spvv8_len);
    for (int t = 0; t < NUM_THREADS; t++) {
        free(s->use_x_indices[t]);
        free(s->ecr_xx_val[t]);
        free(s->tasks[t]);
    }
    free(s->use_x_indices);
    free(s->ecr_xx_val);
    free(s->tasks);
    free(s->reorder_mat.nnz);
    free(s->reorder_mat.col);
    free(s->reorder_mat.row_begin);
    free(s->reorder_mat.row_end);
    free(s->reorder_mat.task_start);
    free(s->reorder_mat.task_end);
    free(s);
}

/* ------------------------------------------------------------------ L */
void *ref_build_l(void *h) { return Build_VNEC_L((SpB_Matrix)h); }

void ref_l_export(void *p, int32_t *coords)
{
    VNEC_L *l = (VNEC_L *)p;
    for (int t = 0; t < NUM_THREADS; t++) {
        coords[4 * t + 0] = l->thread_coord_start[t].x;
        coords[4 * t + 1] = l->thread_coord_start[t].y;
        coords[4 * t + 2] = l->thread_coord_end[t].x;
        coords[4 * t + 3] = l->thread_coord_end[t].y;
    }
}

void ref_spmv_l(void *p, void *h, const double *x, double *y, int reps)
{
    SpB_Matrix A = (SpB_Matrix)h;
    SpB_Vector xv = make_vec(x, (int64_t)A->col), yv = make_vec(NULL, (int64_t)A->row);
    for (int r = 0; r < reps; r++) SpMV_VNEC_L_FP64(yv, A, xv, (VNEC_L *)p);
    memcpy(y, ((SpB_Vector_Dense *)yv)->values, sizeof(double) * A->row);
    SpB_Vector_free(&xv);
    SpB_Vector_free(&yv);
}

void ref_free_l(void *p)
{
    VNEC_L *l = (VNEC_L *)p;
    free(l->diagonal_start);
    free(l->diagonal_end);
    free(l->nz_indices);
    free(l->thread_coord_start);
    free(l->thread_coord_end);
    free(l);
}

/* Timed loop for the CPU baseline: one warm-up then `reps` SpMVs of the
 * chosen variant (0 = D, 1 = S, 2 = L); returns seconds per SpMV measured
 * with a monotonic clock, like the reference harness
 * (test_spmv_vnec.cpp:205-214, which uses gettimeofday). */
#include <time.h>
double ref_time_spmv(int variant, void *p, void *h, const double *x, double *y, int reps)
{
    SpB_Matrix A = (SpB_Matrix)h;
    SpB_Vector xv = make_vec(x, (int64_t)A->col), yv = make_vec(NULL, (int64_t)A->row);
    struct timespec t0, t1;
    for (int w = 0; w < 1; w++) {
        if (variant == 0) SpMV_VNEC_D_FP64(yv, A, xv, (VNEC_D_FP64 *)p);
        else if (variant == 1) SpMV_VNEC_S_FP64(yv, A, xv, (VNEC_S_FP64 *)p);
        else SpMV_VNEC_L_FP64(yv, A, xv, (VNEC_L *)p);
    }
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int r = 0; r < reps; r++) {
        if (variant == 0) SpMV_VNEC_D_FP64(yv, A, xv, (VNEC_D_FP64 *)p);
        else if (variant == 1) SpMV_VNEC_S_FP64(yv, A, xv, (VNEC_S_FP64 *)p);
        else SpMV_VNEC_L_FP64(yv, A, xv, (VNEC_L *)p);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    memcpy(y, ((SpB_Vector_Dense *)yv)->values, sizeof(double) * A->row);
    SpB_Vector_free(&xv);
    SpB_Vector_free(&yv);
    return ((double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec)) / reps;
}

/* ------------------------------------------------------------- FP32 API
 * The reference's FP32 twins (spmv_vnec_D_for_mxv.c:2-241,
 * spmv_vnec_S_for_mxv.c:3-334, spmv_vnec_L_for_mxv.c:36-122, IRD at
 * nLanes_f32 / IRD_thr_fp32) on an SpB_FP32 matrix built from a caller CSR
 * (values padded by 32 floats, as for FP64). */
int ref_lanes_f32(void) { return nLanes_f32; }

void *ref_matrix32(int64_t nrows, int64_t ncols, const int64_t *ptr, const int32_t *col,
                   const float *val)
{
    ensure_init();
    SpB_Matrix A = NULL;
    if (SpB_Matrix_new(&A, SpB_FP32, (SpB_Index)nrows, (SpB_Index)ncols) != SpB_SUCCESS) return NULL;
    SpB_Matrix_CSC_or_CSR *a = (SpB_Matrix_CSC_or_CSR *)A;
    int64_t nnz = ptr[nrows];
    free(a->ptr);
    a->ptr = (SpB_Index *)malloc(sizeof(SpB_Index) * (size_t)(nrows + 1));
    a->indices = (SpB_Index *)malloc(sizeof(SpB_Index) * (size_t)(nnz + 32));
    a->val = calloc((size_t)(nnz + 32), sizeof(float));
    for (int64_t i = 0; i <= nrows; i++) a->ptr[i] = (SpB_Index)ptr[i];
    for (int64_t j = 0; j < nnz; j++) a->indices[j] = (SpB_Index)col[j];
    memcpy(a->val, val, sizeof(float) * (size_t)nnz);
    a->ptr_len = (SpB_Index)(nrows + 1);
    a->nnz = (SpB_Index)nnz;
    a->capacity = (SpB_Index)nnz;
    return A;
}

static SpB_Vector make_vec32(const float *v, int64_t n)
{
    SpB_Vector x = NULL;
    SpB_Vector_new(&x, SpB_FP32, (SpB_Index)n);
    if (v) memcpy(((SpB_Vector_Dense *)x)->values, v, sizeof(float) * (size_t)n);
    return x;
}

int ref_ird32(void *h) { return (int)IRD_VNEC((SpB_Matrix)h, SpB_FP32); }

void *ref_build_d32(void *h) { return Build_VNEC_D_FP32((SpB_Matrix)h); }

int64_t ref_d32_groups(void *p, void *h)
{
    return ((VNEC_D_FP32 *)p)->v_row_ptr[((SpB_Matrix)h)->row];
}

void ref_d32_export(void *p, void *h, int32_t *coords, int32_t *ecr, int32_t *nec, int32_t *win,
                    int32_t *v_row_ptr, int32_t *col_start, float *val_align)
{
    VNEC_D_FP32 *d = (VNEC_D_FP32 *)p;
    SpB_Matrix_CSC_or_CSR *a = (SpB_Matrix_CSC_or_CSR *)h;
    int64_t nrows = (int64_t)a->matrix_base.row, nnz = (int64_t)a->nnz;
    int64_t off = 0;
    for (int t = 0; t < NUM_THREADS; t++) {
        coords[4 * t + 0] = d->thread_coord_start[t].x;
        coords[4 * t + 1] = d->thread_coord_start[t].y;
        coords[4 * t + 2] = d->thread_coord_end[t].x;
        coords[4 * t + 3] = d->thread_coord_end[t].y;
        nec[t] = d->NEC_NUM[t];
        memcpy(win + off, d->use_x_indices[t], sizeof(int32_t) * (size_t)d->NEC_NUM[t]);
        off += d->NEC_NUM[t];
    }
    memcpy(ecr, d->ecr_indices, sizeof(int32_t) * (size_t)nnz);
    memcpy(v_row_ptr, d->v_row_ptr, sizeof(int32_t) * (size_t)(nrows + 1));
    int64_t G = d->v_row_ptr[nrows];
    memcpy(col_start, d->col_start, sizeof(int32_t) * (size_t)G);
    memcpy(val_align, d->val_align, sizeof(float) * (size_t)(G * nLanes_f32));
}

void ref_spmv_d32(void *p, void *h, const float *x, float *y, int reps)
{
    SpB_Matrix A = (SpB_Matrix)h;
    SpB_Vector xv = make_vec32(x, (int64_t)A->col), yv = make_vec32(NULL, (int64_t)A->row);
    for (int r = 0; r < reps; r++) SpMV_VNEC_D_FP32(yv, A, xv, (VNEC_D_FP32 *)p);
    memcpy(y, ((SpB_Vector_Dense *)yv)->values, sizeof(float) * A->row);
    SpB_Vector_free(&xv);
    SpB_Vector_free(&yv);
}

void ref_free_d32(void *p)
{
    VNEC_D_FP32 *d = (VNEC_D_FP32 *)p;
    free(d->thread_coord_start);
    free(d->thread_coord_end);
    free(d->ecr_indices);
    free(d->NEC_NUM);
    for (int t = 0; t < NUM_THREADS; t++) {
        free(d->use_x_indices[t]);
        _mm_free(d->ecr_xx_val[t]);
    }
    free(d->use_x_indices);
    free(d->ecr_xx_val);
    free(d->col_start);
    free(d->v_row_ptr);
    _mm_free(d->val_align);
    free(d);
}

void *ref_build_s32(void *h) { return Build_VNEC_S_FP32((SpB_Matrix)h); }

void ref_s32_export(void *p, void *h, int32_t *nec, int32_t *win, int32_t *spvv_len, int32_t *tasks,
                    int32_t *row_begin, int32_t *row_end, int32_t *s_col, float *s_val,
                    int32_t *task_se)
{
    VNEC_S_FP32 *s = (VNEC_S_FP32 *)p;
    SpB_Matrix_CSC_or_CSR *a = (SpB_Matrix_CSC_or_CSR *)h;
    int64_t nrows = (int64_t)a->matrix_base.row, nnz = (int64_t)a->nnz;
    int64_t off = 0, toff = 0;
    for (int t = 0; t < NUM_THREADS; t++) {
        nec[t] = s->NEC_NUM[t];
        memcpy(win + off, s->use_x_indices[t], sizeof(int32_t) * (size_t)s->NEC_NUM[t]);
        off += s->NEC_NUM[t];
        spvv_len[t] = s->spvv16_len[t];
        int64_t nr = s->reorder_mat.task_end[t] - s->reorder_mat.task_start[t];
        memcpy(tasks + toff, s->tasks[t], sizeof(int32_t) * (size_t)nr);
        toff += nr;
        task_se[2 * t + 0] = s->reorder_mat.task_start[t];
        task_se[2 * t + 1] = s->reorder_mat.task_end[t];
    }
    memcpy(row_begin, s->reorder_mat.row_begin, sizeof(int32_t) * (size_t)nrows);
    memcpy(row_end, s->reorder_mat.row_end, sizeof(int32_t) * (size_t)nrows);
    memcpy(s_col, s->reorder_mat.col, sizeof(int32_t) * (size_t)nnz);
    memcpy(s_val, s->reorder_mat.nnz, sizeof(float) * (size_t)nnz);
}

void ref_spmv_s32(void *p, void *h, const float *x, float *y, int reps)
{
    SpB_Matrix A = (SpB_Matrix)h;
    SpB_Vector xv = make_vec32(x, (int64_t)A->col), yv = make_vec32(NULL, (int64_t)A->row);
    for (int r = 0; r < reps; r++) SpMV_VNEC_S_FP32(yv, A, xv, (VNEC_S_FP32 *)p);
    memcpy(y, ((SpB_Vector_Dense *)yv)->values, sizeof(float) * A->row);
    SpB_Vector_free(&xv);
    SpB_Vector_free(&yv);
}

void ref_free_s32(void *p)
{
    VNEC_S_FP32 *s = (VNEC_S_FP32 *)p;
    free(s->NEC_NUM);
    free(s->spvv16_len);
    for (int t = 0; t < NUM_THREADS; t++) {
        free(s->use_x_indices[t]);
        free(s->ecr_xx_val[t]);
        free(s->tasks[t]);
    }
    free(s->use_x_indices);
    free(s->ecr_xx_val);
    free(s->tasks);
    free(s->reorder_mat.nnz);
    free(s->reorder_mat.col);
    free(s->reorder_mat.row_begin);
    free(s->reorder_mat.row_end);
    free(s->reorder_mat.task_start);
    free(s->reorder_mat.task_end);
    free(s);
}

void ref_spmv_l32(void *p, void *h, const float *x, float *y, int reps)
{
    SpB_Matrix A = (SpB_Matrix)h;
    SpB_Vector xv = make_vec32(x, (int64_t)A->col), yv = make_vec32(NULL, (int64_t)A->row);
    for (int r = 0; r < reps; r++) SpMV_VNEC_L_FP32(yv, A, xv, (VNEC_L *)p);
    memcpy(y, ((SpB_Vector_Dense *)yv)->values, sizeof(float) * A->row);
    SpB_Vector_free(&xv);
    SpB_Vector_free(&yv);
}
