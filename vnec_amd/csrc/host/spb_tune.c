/*
 * spb_tune.c — the exhaustive autotuner, the integer matrix statistics, the
 * learned selector (an FP32 MLP trained and run by the library's own HIP
 * kernels) and the selection modes a plan build can ask for.
 *
 * Candidates: the plan configurations a selector may choose (variant x
 * merge-path granularity x schedule).  The autotuner builds and times each
 * with hipGraph replay on the library stream; its fastest is the label
 * the MLP is trained on and the yardstick of the MLP's regret.
 *
 * The reference's only selector is IRD_VNEC (ird_calc.c:2-123) plus the
 * row-count gate of its harness (test_spmv_vnec.cpp:197-256); the
 * VNEC_SELECT_REFERENCE mode maps that decision onto the GPU variants.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "spb_internal.h"

/* variant, ipt, lanes, encoding, nrhs, grid (0 persistent, -1 one
 * workgroup per tile) */
static const vnec_tuning CANDIDATES[VNEC_NCAND] = {
    {VNEC_VAR_S, 2, 64, 0, 1, -1},  {VNEC_VAR_S, 3, 64, 0, 1, -1},  {VNEC_VAR_S, 5, 64, 0, 1, -1},
    {VNEC_VAR_L, 5, 64, 0, 1, 0},   {VNEC_VAR_L, 7, 64, 0, 1, 0},
    {VNEC_VAR_C, 7, VNEC_LANES_WIDE | VNEC_LANES_GATHER, 0, 1, -1},
    {VNEC_VAR_L, 11, 64, 0, 1, 0},  {VNEC_VAR_L, 5, 64, 0, 1, -1},  {VNEC_VAR_L, 7, 64, 0, 1, -1},
    {VNEC_VAR_L, 11, 64, 0, 1, -1}, {VNEC_VAR_E, 5, 64, 0, 1, -1},  {VNEC_VAR_E, 7, 64, 0, 1, -1},
    {VNEC_VAR_E, 5, 64, 0, 1, 0},   {VNEC_VAR_E, 7, 64, 0, 1, 0},
    {VNEC_VAR_D, 7, 4, 0, 1, -1},   {VNEC_VAR_D, 7, 8, 0, 1, -1},   {VNEC_VAR_D, 7, 16, 0, 1, -1},
    {VNEC_VAR_C, 9, 64, 0, 1, 0},   {VNEC_VAR_C, 11, 64, 0, 1, 0},  {VNEC_VAR_C, 7, 64, 0, 1, -1},
    {VNEC_VAR_C, 11, 64, 0, 1, -1}, {VNEC_VAR_G, 8, 64, 0, 1, -1},  {VNEC_VAR_C, 9, VNEC_LANES_WIDE, 0, 1, -1},
};
/* Candidates named by the fixed selection modes. */
#define CAND_S3 1
#define CAND_L7_PERSISTENT 4
#define CAND_D8 15 /* the reference's own block width */
#define CAND_G 21  /* pattern slices */
/* Pattern slices apply when at least this fraction of the rows fits the
 * matrix's dominant column-offset pattern (the others run as a sub-plan). */
#define G_MIN_ROW_FRACTION 0.5

/* Slice candidates apply only up to this padding (slots / nnz). */
#define TUNE_MAX_PADDING 3.0
/* D candidates apply while their value blocks (at most 8 * lanes bytes per
 * nonzero) fit this many bytes. */
#define TUNE_MAX_D_BYTES (1ull << 32)
/* Row count from which the reference dispatch always takes VNEC-L
 * (test_spmv_vnec.cpp:197). */
#define REF_L_ROWS 3000000
/* The rule selector's slice-padding threshold. */
#define RULE_MAX_PADDING 1.15

int vnec_candidates_ext(vnec_tuning *out, int max)
{
    int n = max < VNEC_NCAND ? max : VNEC_NCAND;
    if (out)
        for (int i = 0; i < n; i++) out[i] = CANDIDATES[i];
    return VNEC_NCAND;
}

uint32_t vnec_candidate_mask_from_ints_ext(const uint64_t v[VNEC_NFEAT_INTS])
{
    uint64_t nnz = v[VF_NNZ], slots = v[VF_SLICE_SLOTS], ns = (v[VF_NROWS] + 63) / 64;
    int slices_ok = nnz > 0 && (double)slots <= TUNE_MAX_PADDING * (double)nnz &&
                    slots + ns < (uint64_t)INT32_MAX;
    uint32_t m = 0;
    for (int c = 0; c < VNEC_NCAND; c++) {
        if (CANDIDATES[c].variant == VNEC_VAR_S && !slices_ok) continue;
        if (CANDIDATES[c].variant == VNEC_VAR_D &&
            (nnz == 0 || nnz * 8ull * (uint64_t)CANDIDATES[c].lanes > TUNE_MAX_D_BYTES))
            continue;
        /* pattern slices: every row holds VNEC_PATTERN_MAX slots at most */
        if (CANDIDATES[c].variant == VNEC_VAR_G &&
            (nnz == 0 || ns * 64ull * VNEC_PATTERN_MAX >= (uint64_t)INT32_MAX))
            continue;
        m |= 1u << c;
    }
    return m;
}

/* ------------------------------------------------------------ statistics */
#define LEN_BINS 65536

/* Sum over pairs i < j of |l_i - l_j| from the length histogram, visiting
 * lengths in increasing order: each row of length v adds v * (rows below)
 * - (their total length).  Wraps modulo 2^64. */
static uint64_t gini_numerator(const uint32_t *hist)
{
    uint64_t below = 0, below_sum = 0, acc = 0;
    for (uint64_t v = 0; v < LEN_BINS; v++) {
        uint64_t c = hist[v];
        if (!c) continue;
        acc += c * (v * below - below_sum);
        below += c;
        below_sum += c * v;
    }
    return acc;
}

/* Distinct columns and runs of consecutive distinct columns, summed over
 * the tiles of a merge-path partition with about `items` items per tile. */
static SpB_Info window_stats(const spb_csr *a, int64_t items, uint64_t *nec_sum, uint64_t *runs_sum)
{
    hipStream_t s = vnec_stream();
    int64_t nr = (int64_t)a->base.row, nnz = (int64_t)a->nnz;
    *nec_sum = *runs_sum = 0;
    if (nnz == 0) return SpB_SUCCESS;
    int64_t M = nr + nnz, P = (M + items - 1) / items;
    int32_t *co = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * 2 * (size_t)(P + 1), NULL);
    int32_t *uniq = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)nnz, NULL);
    int32_t *nec = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)P, NULL);
    int32_t *nrun = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)P, NULL);
    int32_t *hn = (int32_t *)malloc(sizeof(int32_t) * (size_t)P);
    int32_t *hr = (int32_t *)malloc(sizeof(int32_t) * (size_t)P);
    SpB_Info rc = (co && uniq && nec && nrun && hn && hr) ? SpB_SUCCESS : vnec_fail("out of memory");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_partition(a->d_ptr, nr, nnz, P, 0, co, s), "partition");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_ecr_tiles_pass1(a->d_col, co, P, NULL, NULL, uniq, nec, nrun, s),
                         "window columns");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(hn, nec, sizeof(int32_t) * (size_t)P, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(hr, nrun, sizeof(int32_t) * (size_t)P, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    if (rc == SpB_SUCCESS)
        for (int64_t t = 0; t < P; t++) {
            *nec_sum += (uint64_t)hn[t];
            *runs_sum += (uint64_t)hr[t];
        }
    vnec_free_dev(co);
    vnec_free_dev(uniq);
    vnec_free_dev(nec);
    vnec_free_dev(nrun);
    free(hn);
    free(hr);
    return rc;
}

SpB_Info vnec_feature_ints_ext(const SpB_Matrix A, uint64_t ints[VNEC_NFEAT_INTS])
{
    spb_csr *a = spb_as_csr(A);
    if (!a || !ints) return vnec_fail("vnec_feature_ints_ext: bad arguments");
    if (!vnec_runtime_ready() || spb_matrix_upload(a) != SpB_SUCCESS) return SpB_FAILURE;
    if (a->feat_ready) { /* computed since the matrix last changed */
        memcpy(ints, a->feat_ints, sizeof a->feat_ints);
        return SpB_SUCCESS;
    }
    hipStream_t s = vnec_stream();
    int64_t nr = (int64_t)a->base.row, nc = (int64_t)a->base.col;
    uint64_t *d = (uint64_t *)vnec_malloc_dev(sizeof(uint64_t) * VNEC_NFEAT_INTS, NULL);
    uint32_t *dh = (uint32_t *)vnec_malloc_dev(sizeof(uint32_t) * LEN_BINS, NULL);
    uint32_t *hh = (uint32_t *)malloc(sizeof(uint32_t) * LEN_BINS);
    uint64_t h[VNEC_NFEAT_INTS];
    SpB_Info rc = (d && dh && hh) ? SpB_SUCCESS : vnec_fail("out of memory");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemsetAsync(d, 0, sizeof(uint64_t) * VNEC_NFEAT_INTS, s), "memset");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemsetAsync(dh, 0, sizeof(uint32_t) * LEN_BINS, s), "memset");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(vnec_hip_feature_rows(a->d_ptr, nr, d, dh, s), "row statistics");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_feature_cols(a->d_ptr, a->d_col, nr, nc, d, s), "column statistics");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(hh, dh, sizeof(uint32_t) * LEN_BINS, hipMemcpyDeviceToHost, s), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    if (rc == SpB_SUCCESS) {
        memcpy(ints, h, sizeof h);
        ints[VF_NROWS] = (uint64_t)nr;
        ints[VF_NCOLS] = (uint64_t)nc;
        ints[VF_NNZ] = (uint64_t)a->nnz;
        ints[VF_GINI_NUM] = gini_numerator(hh);
        rc = window_stats(a, 4096, &ints[VF_NEC_SMALL], &ints[VF_RUNS_SMALL]);
    }
    if (rc == SpB_SUCCESS) rc = window_stats(a, 16384, &ints[VF_NEC_LARGE], &ints[VF_RUNS_LARGE]);
    vnec_free_dev(d);
    vnec_free_dev(dh);
    free(hh);
    if (rc == SpB_SUCCESS) {
        memcpy(a->feat_ints, ints, sizeof a->feat_ints);
        a->feat_ready = 1;
    }
    return rc;
}

void vnec_features_from_ints_ext(const uint64_t v[VNEC_NFEAT_INTS], float f[VNEC_NFEAT])
{
    double n = (double)v[VF_NROWS], m = (double)v[VF_NCOLS], z = (double)v[VF_NNZ];
    double nn = n > 0 ? n : 1.0, zz = z > 0 ? z : 1.0, mm = m > 0 ? m : 1.0;
    double mean = z / nn;
    double ex2 = (double)v[VF_SUM_LEN2] / nn, ex3 = (double)v[VF_SUM_LEN3] / nn;
    double var = ex2 - mean * mean;
    double sd = var > 0 ? sqrt(var) : 0.0;
    double skew = sd > 0 ? (ex3 - 3.0 * mean * var - mean * mean * mean) / (sd * sd * sd) : 0.0;
    skew = skew > 50.0 ? 50.0 : (skew < -50.0 ? -50.0 : skew);
    int k = 0;
    f[k++] = (float)log2(n + 1.0);
    f[k++] = (float)log2(m + 1.0);
    f[k++] = (float)log2(z + 1.0);
    f[k++] = (float)log2(mean + 1.0);
    f[k++] = (float)(mean > 0 ? sd / mean : 0.0);       /* coefficient of variation */
    f[k++] = (float)(skew / 10.0);                      /* skewness of row lengths  */
    f[k++] = (float)log2((double)v[VF_MAX_LEN] + 1.0);
    f[k++] = (float)((double)v[VF_EMPTY_ROWS] / nn);
    for (int b = 0; b < VNEC_FEAT_NBUCKETS; b++) f[k++] = (float)((double)v[VF_HIST + b] / nn);
    f[k++] = (float)((double)v[VF_GINI_NUM] / (nn * zz)); /* Gini of row lengths      */
    f[k++] = (float)((double)v[VF_SUM_DIST] / zz / mm);  /* mean distance / ncols    */
    f[k++] = (float)((double)v[VF_MAX_DIST] / mm);       /* bandwidth / ncols        */
    f[k++] = (float)((double)v[VF_NEAR_DIAG] / zz);
    f[k++] = (float)((double)v[VF_NEC_SMALL] / zz);      /* column reuse, 4096 items */
    f[k++] = (float)((double)v[VF_RUNS_SMALL] / zz);
    f[k++] = (float)((double)v[VF_NEC_LARGE] / zz);      /* column reuse, 16384 items */
    f[k++] = (float)log2((double)v[VF_SLICE_SLOTS] / zz + 1.0);
}

SpB_Info vnec_candidate_mask_ext(const SpB_Matrix A, uint32_t *mask)
{
    uint64_t ints[VNEC_NFEAT_INTS];
    if (!mask) return vnec_fail("vnec_candidate_mask_ext: NULL output");
    if (vnec_feature_ints_ext(A, ints) != SpB_SUCCESS) return SpB_FAILURE;
    *mask = vnec_candidate_mask_from_ints_ext(ints);
    if ((*mask >> CAND_G) & 1u) {
        /* pattern slices need most rows on one column-offset pattern */
        spb_csr *a = spb_as_csr(A);
        if (!a->pattern_ready) {
            if (vnec_pattern_coverage(a, &a->pattern_rows) != SpB_SUCCESS) return SpB_FAILURE;
            a->pattern_ready = 1;
        }
        if ((double)a->pattern_rows < G_MIN_ROW_FRACTION * (double)a->base.row) *mask &= ~(1u << CAND_G);
    }
    return SpB_SUCCESS;
}

/* -------------------------------------------------------------- autotuner */
SpB_Info vnec_autotune_ext(const SpB_Matrix A, int reps, double times_ms[VNEC_NCAND], int *best)
{
    spb_csr *a = spb_as_csr(A);
    if (!a || !times_ms || reps < 1) return vnec_fail("vnec_autotune_ext: bad arguments");
    if (a->base.element_type != SpB_FP64)
        return vnec_fail("vnec_autotune_ext: the candidates are FP64 kernels; A must hold FP64 values");
    uint32_t mask = 0;
    if (vnec_candidate_mask_ext(A, &mask) != SpB_SUCCESS) return SpB_FAILURE;
    int64_t nr = (int64_t)a->base.row, nc = (int64_t)a->base.col;
    double *hx = (double *)malloc(sizeof(double) * (size_t)(nc ? nc : 1));
    double *dx = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(nc ? nc : 1), NULL);
    double *dy = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(nr ? nr : 1), NULL);
    SpB_Info rc = (hx && dx && dy) ? SpB_SUCCESS : vnec_fail("out of memory");
    for (int64_t i = 0; rc == SpB_SUCCESS && i < nc; i++) hx[i] = 1.0 / (double)(1 + (i % 7));
    if (rc == SpB_SUCCESS && nc)
        rc = vnec_hip_ok(hipMemcpy(dx, hx, sizeof(double) * (size_t)nc, hipMemcpyHostToDevice), "h2d");
    int bi = -1;
    for (int c = 0; c < VNEC_NCAND; c++) times_ms[c] = INFINITY;
    for (int c = 0; rc == SpB_SUCCESS && c < VNEC_NCAND; c++) {
        if (!((mask >> c) & 1u)) continue;
        vnec_plan p = NULL;
        if (vnec_plan_build_ext(&p, A, &CANDIDATES[c]) != SpB_SUCCESS) continue; /* stays INFINITY */
        double tbest = INFINITY;
        for (int r = 0; rc == SpB_SUCCESS && r < 3; r++) {
            double ms = 0.0;
            rc = vnec_time_dev_ext(p, dx, dy, reps, 1, 0, &ms);
            if (rc == SpB_SUCCESS && ms < tbest) tbest = ms;
        }
        if (p) vnec_plan_free_ext(&p);
        times_ms[c] = tbest;
        if (rc == SpB_SUCCESS && (bi < 0 || tbest < times_ms[bi])) bi = c;
    }
    free(hx);
    vnec_free_dev(dx);
    vnec_free_dev(dy);
    if (rc != SpB_SUCCESS) return rc;
    if (bi < 0) return vnec_fail("vnec_autotune_ext: no candidate applies");
    a->tuned_cand = bi;
    if (best) *best = bi;
    return SpB_SUCCESS;
}

/* ------------------------------------------------------------------ MLP */
#define NPAR VNEC_MLP_NPARAMS
#define GRAD_CHUNK 16

typedef struct {
    int installed; /* 1: g_params etc. were set by vnec_selector_install_ext */
    float params[NPAR], mu[VNEC_NFEAT], sd[VNEC_NFEAT];
    float *d_params, *d_mu, *d_sd; /* device copies of the model in force */
    int dev_valid;
} selector_state;
static selector_state g_sel;

static int model_trained(void) { return g_sel.installed || vnec_model_trained; }

static SpB_Info upload(void **dst, const void *src, size_t bytes)
{
    *dst = vnec_malloc_dev(bytes ? bytes : 1, NULL);
    if (!*dst) return SpB_FAILURE;
    if (!bytes) return SpB_SUCCESS;
    return vnec_hip_ok(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, vnec_stream()), "h2d");
}

static SpB_Info download(void *dst, const void *src, size_t bytes)
{
    SpB_Info rc = vnec_hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, vnec_stream()), "d2h");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(vnec_stream()), "sync");
    return rc;
}

SpB_Info vnec_mlp_forward_ext(const float *params, const float *X, int64_t B, float *out)
{
    if (!params || !X || !out || B < 0) return vnec_fail("vnec_mlp_forward_ext: bad arguments");
    if (!vnec_runtime_ready()) return SpB_FAILURE;
    void *dp = NULL, *dx = NULL;
    float *dout = (float *)vnec_malloc_dev(sizeof(float) * (size_t)(B * VNEC_NCAND + 1), NULL);
    SpB_Info rc = dout ? upload(&dp, params, sizeof(float) * NPAR) : SpB_FAILURE;
    if (rc == SpB_SUCCESS) rc = upload(&dx, X, sizeof(float) * (size_t)(B * VNEC_NFEAT));
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_mlp_forward((const float *)dp, (const float *)dx, B, dout, vnec_stream()),
                         "mlp forward");
    if (rc == SpB_SUCCESS) rc = download(out, dout, sizeof(float) * (size_t)(B * VNEC_NCAND));
    vnec_free_dev(dp);
    vnec_free_dev(dx);
    vnec_free_dev(dout);
    return rc;
}

SpB_Info vnec_mlp_grad_ext(const float *params, const float *X, const float *T, const float *mask,
                           int64_t B, float *grad, float *loss)
{
    if (!params || !X || !T || !mask || !grad || B <= 0)
        return vnec_fail("vnec_mlp_grad_ext: bad arguments");
    if (!vnec_runtime_ready()) return SpB_FAILURE;
    double den = 0.0;
    for (int64_t i = 0; i < B * VNEC_NCAND; i++) den += mask[i];
    if (den <= 0.0) return vnec_fail("vnec_mlp_grad_ext: empty mask");
    int64_t nch = (B + GRAD_CHUNK - 1) / GRAD_CHUNK;
    void *dp = NULL, *dx = NULL, *dt = NULL, *dm = NULL;
    float *part = (float *)vnec_malloc_dev(sizeof(float) * (size_t)(nch * NPAR), NULL);
    float *lpart = (float *)vnec_malloc_dev(sizeof(float) * (size_t)nch, NULL);
    float *dg = (float *)vnec_malloc_dev(sizeof(float) * NPAR, NULL);
    float *hl = (float *)malloc(sizeof(float) * (size_t)nch);
    SpB_Info rc = (part && lpart && dg && hl) ? SpB_SUCCESS : vnec_fail("out of memory");
    if (rc == SpB_SUCCESS) rc = upload(&dp, params, sizeof(float) * NPAR);
    if (rc == SpB_SUCCESS) rc = upload(&dx, X, sizeof(float) * (size_t)(B * VNEC_NFEAT));
    if (rc == SpB_SUCCESS) rc = upload(&dt, T, sizeof(float) * (size_t)(B * VNEC_NCAND));
    if (rc == SpB_SUCCESS) rc = upload(&dm, mask, sizeof(float) * (size_t)(B * VNEC_NCAND));
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_mlp_grad((const float *)dp, (const float *)dx, (const float *)dt,
                                           (const float *)dm, NULL, B, GRAD_CHUNK, (float)(1.0 / den),
                                           part, lpart, vnec_stream()), "mlp gradient");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_mlp_reduce_adam(part, (int)nch, dg, (float *)dp, NULL, NULL, 0.f, 0.f,
                                                  1.f, 1.f, vnec_stream()), "gradient reduce");
    if (rc == SpB_SUCCESS) rc = download(grad, dg, sizeof(float) * NPAR);
    if (rc == SpB_SUCCESS) rc = download(hl, lpart, sizeof(float) * (size_t)nch);
    if (rc == SpB_SUCCESS && loss) {
        float l = 0.f;
        for (int64_t c = 0; c < nch; c++) l += hl[c];
        *loss = l;
    }
    vnec_free_dev(dp);
    vnec_free_dev(dx);
    vnec_free_dev(dt);
    vnec_free_dev(dm);
    vnec_free_dev(part);
    vnec_free_dev(lpart);
    vnec_free_dev(dg);
    free(hl);
    return rc;
}

/* splitmix64: the shuffle order of the training samples. */
static uint64_t mix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

SpB_Info vnec_mlp_train_ext(float *params, const float *X, const float *T, const float *mask,
                            int64_t N, int epochs, int batch, float lr, float weight_decay,
                            uint64_t seed, float *loss_hist)
{
    if (!params || !X || !T || !mask || N <= 0 || epochs < 0 || batch <= 0 || lr <= 0.f)
        return vnec_fail("vnec_mlp_train_ext: bad arguments");
    if (N > INT32_MAX) return vnec_fail("vnec_mlp_train_ext: too many samples");
    if (!vnec_runtime_ready()) return SpB_FAILURE;
    hipStream_t s = vnec_stream();
    int64_t maxch = (batch + GRAD_CHUNK - 1) / GRAD_CHUNK;
    int64_t nbatches = (N + batch - 1) / batch;
    void *dp = NULL, *dx = NULL, *dt = NULL, *dm = NULL;
    float *part = (float *)vnec_malloc_dev(sizeof(float) * (size_t)(maxch * NPAR), NULL);
    float *lpart = (float *)vnec_malloc_dev(sizeof(float) * (size_t)(nbatches * maxch), NULL);
    float *am = (float *)vnec_malloc_dev(sizeof(float) * NPAR, NULL);
    float *av = (float *)vnec_malloc_dev(sizeof(float) * NPAR, NULL);
    int32_t *didx = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)N, NULL);
    int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    double *rowmask = (double *)malloc(sizeof(double) * (size_t)N);
    float *hl = (float *)malloc(sizeof(float) * (size_t)(nbatches * maxch));
    SpB_Info rc = (part && lpart && am && av && didx && perm && rowmask && hl) ? SpB_SUCCESS
                                                                              : vnec_fail("out of memory");
    if (rc == SpB_SUCCESS) rc = upload(&dp, params, sizeof(float) * NPAR);
    if (rc == SpB_SUCCESS) rc = upload(&dx, X, sizeof(float) * (size_t)(N * VNEC_NFEAT));
    if (rc == SpB_SUCCESS) rc = upload(&dt, T, sizeof(float) * (size_t)(N * VNEC_NCAND));
    if (rc == SpB_SUCCESS) rc = upload(&dm, mask, sizeof(float) * (size_t)(N * VNEC_NCAND));
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemsetAsync(am, 0, sizeof(float) * NPAR, s), "memset");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemsetAsync(av, 0, sizeof(float) * NPAR, s), "memset");
    for (int64_t i = 0; rc == SpB_SUCCESS && i < N; i++) {
        double r = 0.0;
        for (int c = 0; c < VNEC_NCAND; c++) r += mask[i * VNEC_NCAND + c];
        rowmask[i] = r;
        perm[i] = (int32_t)i;
    }
    uint64_t rng = seed;
    int64_t step = 0;
    for (int e = 0; rc == SpB_SUCCESS && e < epochs; e++) {
        for (int64_t i = N - 1; i > 0; i--) { /* Fisher-Yates */
            int64_t j = (int64_t)(mix64(&rng) % (uint64_t)(i + 1));
            int32_t t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
        /* the previous epoch's kernels may still read didx */
        rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(hipMemsetAsync(lpart, 0, sizeof(float) * (size_t)(nbatches * maxch), s), "memset");
        if (rc == SpB_SUCCESS)
            rc = vnec_hip_ok(hipMemcpy(didx, perm, sizeof(int32_t) * (size_t)N, hipMemcpyHostToDevice), "h2d");
        for (int64_t b = 0; rc == SpB_SUCCESS && b < nbatches; b++) {
            int64_t b0 = b * batch, nb = b0 + batch <= N ? batch : N - b0;
            double den = 0.0;
            for (int64_t i = 0; i < nb; i++) den += rowmask[perm[b0 + i]];
            if (den <= 0.0) continue;
            int64_t nch = (nb + GRAD_CHUNK - 1) / GRAD_CHUNK;
            step++;
            float bc1 = (float)(1.0 - pow(0.9, (double)step));
            float bc2 = (float)(1.0 - pow(0.999, (double)step));
            rc = vnec_hip_ok(vnec_hip_mlp_grad((const float *)dp, (const float *)dx, (const float *)dt,
                                               (const float *)dm, didx + b0, nb, GRAD_CHUNK,
                                               (float)(1.0 / den), part, lpart + b * maxch, s),
                             "mlp gradient");
            if (rc == SpB_SUCCESS)
                rc = vnec_hip_ok(vnec_hip_mlp_reduce_adam(part, (int)nch, NULL, (float *)dp, am, av, lr,
                                                          weight_decay, bc1, bc2, s), "adam");
        }
        if (rc == SpB_SUCCESS && loss_hist) {
            rc = download(hl, lpart, sizeof(float) * (size_t)(nbatches * maxch));
            /* mean of the minibatch losses (each already normalised) */
            double sum = 0.0;
            int64_t used = 0;
            for (int64_t b = 0; rc == SpB_SUCCESS && b < nbatches; b++) {
                int64_t b0 = b * batch, nb = b0 + batch <= N ? batch : N - b0;
                int64_t nch = (nb + GRAD_CHUNK - 1) / GRAD_CHUNK;
                double l = 0.0;
                for (int64_t c = 0; c < nch; c++) l += hl[b * maxch + c];
                sum += l;
                used++;
            }
            loss_hist[e] = (float)(sum / (double)(used ? used : 1));
        }
    }
    if (rc == SpB_SUCCESS) rc = download(params, dp, sizeof(float) * NPAR);
    vnec_free_dev(dp);
    vnec_free_dev(dx);
    vnec_free_dev(dt);
    vnec_free_dev(dm);
    vnec_free_dev(part);
    vnec_free_dev(lpart);
    vnec_free_dev(am);
    vnec_free_dev(av);
    vnec_free_dev(didx);
    free(perm);
    free(rowmask);
    free(hl);
    return rc;
}

SpB_Info vnec_mlp_select_ext(const float *params, const float *mu, const float *sd,
                             const float *features, const uint32_t *mask, int64_t B, int32_t *choice)
{
    if (!params || !mu || !sd || !features || !mask || !choice || B < 0)
        return vnec_fail("vnec_mlp_select_ext: bad arguments");
    if (!vnec_runtime_ready()) return SpB_FAILURE;
    if (B == 0) return SpB_SUCCESS;
    void *dp = NULL, *dmu = NULL, *dsd = NULL, *df = NULL, *dmk = NULL;
    int32_t *dc = (int32_t *)vnec_malloc_dev(sizeof(int32_t) * (size_t)B, NULL);
    SpB_Info rc = dc ? upload(&dp, params, sizeof(float) * NPAR) : SpB_FAILURE;
    if (rc == SpB_SUCCESS) rc = upload(&dmu, mu, sizeof(float) * VNEC_NFEAT);
    if (rc == SpB_SUCCESS) rc = upload(&dsd, sd, sizeof(float) * VNEC_NFEAT);
    if (rc == SpB_SUCCESS) rc = upload(&df, features, sizeof(float) * (size_t)(B * VNEC_NFEAT));
    if (rc == SpB_SUCCESS) rc = upload(&dmk, mask, sizeof(uint32_t) * (size_t)B);
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_mlp_select((const float *)dp, (const float *)dmu, (const float *)dsd,
                                             (const float *)df, (const uint32_t *)dmk, B, dc, vnec_stream()),
                         "mlp select");
    if (rc == SpB_SUCCESS) rc = download(choice, dc, sizeof(int32_t) * (size_t)B);
    vnec_free_dev(dp);
    vnec_free_dev(dmu);
    vnec_free_dev(dsd);
    vnec_free_dev(df);
    vnec_free_dev(dmk);
    vnec_free_dev(dc);
    return rc;
}

SpB_Info vnec_selector_install_ext(const float *params, const float *mu, const float *sd)
{
    if (params && (!mu || !sd)) return vnec_fail("vnec_selector_install_ext: mu and sd required");
    if (params) {
        for (int k = 0; k < VNEC_NFEAT; k++)
            if (!(sd[k] > 0.f)) return vnec_fail("vnec_selector_install_ext: sd[%d] must be > 0", k);
        memcpy(g_sel.params, params, sizeof g_sel.params);
        memcpy(g_sel.mu, mu, sizeof g_sel.mu);
        memcpy(g_sel.sd, sd, sizeof g_sel.sd);
    }
    g_sel.installed = params != NULL;
    g_sel.dev_valid = 0;
    return SpB_SUCCESS;
}

SpB_Info vnec_selector_info_ext(int *trained, int64_t *train_samples)
{
    if (trained) *trained = model_trained();
    if (train_samples) *train_samples = g_sel.installed ? -1 : vnec_model_train_samples;
    return SpB_SUCCESS;
}

/* Device copy of the model in force (uploaded once per change). */
static SpB_Info selector_device_model(void)
{
    if (g_sel.dev_valid) return SpB_SUCCESS;
    const float *p = g_sel.installed ? g_sel.params : vnec_model_params;
    const float *mu = g_sel.installed ? g_sel.mu : vnec_model_mu;
    const float *sd = g_sel.installed ? g_sel.sd : vnec_model_sd;
    if (!g_sel.d_params) {
        g_sel.d_params = (float *)vnec_malloc_dev(sizeof(float) * NPAR, NULL);
        g_sel.d_mu = (float *)vnec_malloc_dev(sizeof(float) * VNEC_NFEAT, NULL);
        g_sel.d_sd = (float *)vnec_malloc_dev(sizeof(float) * VNEC_NFEAT, NULL);
        if (!g_sel.d_params || !g_sel.d_mu || !g_sel.d_sd) return SpB_FAILURE;
    }
    hipStream_t s = vnec_stream();
    SpB_Info rc = vnec_hip_ok(hipMemcpyAsync(g_sel.d_params, p, sizeof(float) * NPAR, hipMemcpyHostToDevice, s), "h2d");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(g_sel.d_mu, mu, sizeof(float) * VNEC_NFEAT, hipMemcpyHostToDevice, s), "h2d");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(hipMemcpyAsync(g_sel.d_sd, sd, sizeof(float) * VNEC_NFEAT, hipMemcpyHostToDevice, s), "h2d");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipStreamSynchronize(s), "sync");
    g_sel.dev_valid = rc == SpB_SUCCESS;
    return rc;
}

/* The MLP's pick among the applicable candidates that `allow` keeps. */
static SpB_Info select_learned(spb_csr *a, uint32_t allow, int *cand)
{
    uint64_t ints[VNEC_NFEAT_INTS];
    float f[VNEC_NFEAT];
    uint32_t mask = 0;
    if (vnec_feature_ints_ext((SpB_Matrix)a, ints) != SpB_SUCCESS) return SpB_FAILURE;
    vnec_features_from_ints_ext(ints, f);
    if (vnec_candidate_mask_ext((SpB_Matrix)a, &mask) != SpB_SUCCESS) return SpB_FAILURE;
    mask &= allow;
    if (!mask) return vnec_fail("selector: no candidate applies");
    if (selector_device_model() != SpB_SUCCESS) return SpB_FAILURE;
    hipStream_t s = vnec_stream();
    float *df = (float *)vnec_malloc_dev(sizeof f + sizeof(uint32_t) + sizeof(int32_t), NULL);
    if (!df) return SpB_FAILURE;
    uint32_t *dmk = (uint32_t *)(df + VNEC_NFEAT);
    int32_t *dc = (int32_t *)(dmk + 1);
    int32_t c = -1;
    SpB_Info rc = vnec_hip_ok(hipMemcpyAsync(df, f, sizeof f, hipMemcpyHostToDevice, s), "h2d");
    if (rc == SpB_SUCCESS) rc = vnec_hip_ok(hipMemcpyAsync(dmk, &mask, sizeof mask, hipMemcpyHostToDevice, s), "h2d");
    if (rc == SpB_SUCCESS)
        rc = vnec_hip_ok(vnec_hip_mlp_select(g_sel.d_params, g_sel.d_mu, g_sel.d_sd, df, dmk, 1, dc, s),
                         "mlp select");
    if (rc == SpB_SUCCESS) rc = download(&c, dc, sizeof c);
    vnec_free_dev(df);
    if (rc != SpB_SUCCESS) return rc;
    if (c < 0 || c >= VNEC_NCAND) return vnec_fail("selector returned no candidate");
    *cand = c;
    return SpB_SUCCESS;
}

SpB_Info vnec_select_ext(const SpB_Matrix A, int mode, int *candidate)
{
    spb_csr *a = spb_as_csr(A);
    if (!a || !candidate) return vnec_fail("vnec_select_ext: bad arguments");
    if (!vnec_runtime_ready() || spb_matrix_upload(a) != SpB_SUCCESS) return SpB_FAILURE;
    if (mode == VNEC_SELECT_LEARNED && !model_trained()) mode = VNEC_SELECT_RULE;
    switch (mode) {
    case VNEC_SELECT_LEARNED:
        return select_learned(a, ~0u, candidate);
    case VNEC_SELECT_AUTOTUNE: {
        if (a->tuned_cand >= 0) {
            *candidate = a->tuned_cand;
            return SpB_SUCCESS;
        }
        double t[VNEC_NCAND];
        return vnec_autotune_ext(A, 20, t, candidate);
    }
    case VNEC_SELECT_RULE:
        *candidate = a->nnz > 0 && vnec_slice_padding(a) <= RULE_MAX_PADDING ? CAND_S3 : CAND_L7_PERSISTENT;
        return SpB_SUCCESS;
    case VNEC_SELECT_REFERENCE: {
        if ((int64_t)a->base.row >= REF_L_ROWS) {
            *candidate = CAND_L7_PERSISTENT;
            return SpB_SUCCESS;
        }
        int dec = SpB_VNEC_S;
        if (vnec_ird_ext(A, 36, 8, 0.477, NULL, &dec) != SpB_SUCCESS) return SpB_FAILURE;
        if (dec == SpB_VNEC_D) {
            *candidate = CAND_D8;
        } else {
            uint64_t ints[VNEC_NFEAT_INTS];
            if (vnec_feature_ints_ext(A, ints) != SpB_SUCCESS) return SpB_FAILURE;
            uint32_t mask = vnec_candidate_mask_from_ints_ext(ints);
            *candidate = (mask >> CAND_S3) & 1u ? CAND_S3 : CAND_L7_PERSISTENT;
        }
        return SpB_SUCCESS;
    }
    default:
        return vnec_fail("unknown selection mode %d", mode);
    }
}

/* Several right-hand sides run on the merge-path kernel with one workgroup
 * per tile; the autotuner picks its granularity among the compiled ones. */
static const int MULTI_IPT[] = {3, 5, 7};

static SpB_Info autotune_multi(spb_csr *a, int nrhs, int *ipt)
{
    int64_t nr = (int64_t)a->base.row, nc = (int64_t)a->base.col;
    size_t xb = sizeof(double) * (size_t)(nc * nrhs > 0 ? nc * nrhs : 1);
    double *hx = (double *)malloc(xb);
    double *dx = (double *)vnec_malloc_dev(xb, NULL);
    double *dy = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(nr * nrhs > 0 ? nr * nrhs : 1), NULL);
    SpB_Info rc = (hx && dx && dy) ? SpB_SUCCESS : vnec_fail("out of memory");
    for (int64_t i = 0; rc == SpB_SUCCESS && i < nc * nrhs; i++) hx[i] = 1.0 / (double)(1 + (i % 7));
    if (rc == SpB_SUCCESS && nc)
        rc = vnec_hip_ok(hipMemcpy(dx, hx, xb, hipMemcpyHostToDevice), "h2d");
    double best = INFINITY;
    *ipt = 5;
    for (size_t c = 0; rc == SpB_SUCCESS && c < sizeof MULTI_IPT / sizeof MULTI_IPT[0]; c++) {
        vnec_tuning t = {VNEC_VAR_L, MULTI_IPT[c], 64, VNEC_ENC_FP64, nrhs, -1};
        vnec_plan p = NULL;
        if (vnec_plan_build_ext(&p, (SpB_Matrix)a, &t) != SpB_SUCCESS) continue;
        double tb = INFINITY;
        for (int r = 0; rc == SpB_SUCCESS && r < 3; r++) {
            double ms = 0.0;
            rc = vnec_time_dev_ext(p, dx, dy, 10, 1, 0, &ms);
            if (rc == SpB_SUCCESS && ms < tb) tb = ms;
        }
        vnec_plan_free_ext(&p);
        if (tb < best) {
            best = tb;
            *ipt = MULTI_IPT[c];
        }
    }
    free(hx);
    vnec_free_dev(dx);
    vnec_free_dev(dy);
    return rc;
}

/* Best of three hipGraph-timed means of `iters` SpMVs of plan tuning u on
 * a (INFINITY when u does not build). */
static SpB_Info time_tuning(spb_csr *a, const vnec_tuning *u, const void *dx, void *dy, int iters,
                            double *ms_out)
{
    *ms_out = INFINITY;
    vnec_plan p = NULL;
    if (vnec_plan_build_ext(&p, (SpB_Matrix)a, u) != SpB_SUCCESS) {
        vnec_clear_error(); /* not a contender */
        return SpB_SUCCESS;
    }
    SpB_Info rc = SpB_SUCCESS;
    for (int r = 0; rc == SpB_SUCCESS && r < 3; r++) {
        double ms = 0.0;
        rc = vnec_time_dev_ext(p, dx, dy, iters, 1, 0, &ms);
        if (rc == SpB_SUCCESS && ms < *ms_out) *ms_out = ms;
    }
    vnec_plan_free_ext(&p);
    return rc;
}

/* FP32 plan shapes the autotune mode times on an FP32 matrix (the FP64
 * candidates' times do not carry over: values and x are half as wide). */
static const vnec_tuning FP32_SHAPES[] = {
    {VNEC_VAR_S, 2, 64, VNEC_ENC_FP32, 1, -1},  {VNEC_VAR_S, 3, 64, VNEC_ENC_FP32, 1, -1},
    {VNEC_VAR_S, 5, 64, VNEC_ENC_FP32, 1, -1},  {VNEC_VAR_G, 8, 64, VNEC_ENC_FP32, 1, -1},
    {VNEC_VAR_L, 5, 64, VNEC_ENC_FP32, 1, -1},  {VNEC_VAR_L, 7, 64, VNEC_ENC_FP32, 1, -1},
    {VNEC_VAR_L, 11, 64, VNEC_ENC_FP32, 1, -1}, {VNEC_VAR_C, 5, 64, VNEC_ENC_FP32, 1, -1},
    {VNEC_VAR_C, 7, 64, VNEC_ENC_FP32, 1, -1},  {VNEC_VAR_C, 9, 64, VNEC_ENC_FP32, 1, -1},
    {VNEC_VAR_C, 11, 64, VNEC_ENC_FP32, 1, -1}, {VNEC_VAR_D, 7, 16, VNEC_ENC_FP32, 1, -1},
    {VNEC_VAR_C, 9, VNEC_LANES_WIDE, VNEC_ENC_FP32, 1, -1},
    {VNEC_VAR_C, 11, VNEC_LANES_WIDE, VNEC_ENC_FP32, 1, -1},
    {VNEC_VAR_C, 5, VNEC_LANES_WIDE | VNEC_LANES_GATHER, VNEC_ENC_FP32, 1, -1},
    {VNEC_VAR_C, 7, VNEC_LANES_WIDE | VNEC_LANES_GATHER, VNEC_ENC_FP32, 1, -1},
};
#define N_FP32_SHAPES ((int)(sizeof FP32_SHAPES / sizeof FP32_SHAPES[0]))

SpB_Info vnec_autotune_fp32(spb_csr *a, vnec_tuning *t)
{
    if (a->tuned_fp32 >= 0) {
        *t = FP32_SHAPES[a->tuned_fp32];
        return SpB_SUCCESS;
    }
    uint32_t mask = 0;
    if (vnec_candidate_mask_ext((SpB_Matrix)a, &mask) != SpB_SUCCESS) return SpB_FAILURE;
    const int64_t nr = (int64_t)a->base.row, nc = (int64_t)a->base.col, nnz = (int64_t)a->nnz;
    /* the structural conditions of the FP64 candidates: slices (padding),
     * pattern slices (the dominant pattern); D blocks of 4-byte values */
    const int slices = (mask & 7u) != 0, pattern = (mask >> CAND_G) & 1u;
    const int dblocks = nnz > 0 && (uint64_t)nnz * 4ull * 16ull <= TUNE_MAX_D_BYTES;
    float *hx = (float *)malloc(sizeof(float) * (size_t)(nc ? nc : 1));
    float *dx = (float *)vnec_malloc_dev(sizeof(float) * (size_t)(nc ? nc : 1), NULL);
    float *dy = (float *)vnec_malloc_dev(sizeof(float) * (size_t)(nr ? nr : 1), NULL);
    SpB_Info rc = (hx && dx && dy) ? SpB_SUCCESS : vnec_fail("out of memory");
    for (int64_t i = 0; rc == SpB_SUCCESS && i < nc; i++) hx[i] = 1.0f / (float)(1 + (i % 7));
    if (rc == SpB_SUCCESS && nc)
        rc = vnec_hip_ok(hipMemcpy(dx, hx, sizeof(float) * (size_t)nc, hipMemcpyHostToDevice), "h2d");
    double best = INFINITY;
    int bi = 5; /* L7, which applies to every matrix */
    for (int c = 0; rc == SpB_SUCCESS && c < N_FP32_SHAPES; c++) {
        const int v = FP32_SHAPES[c].variant;
        if ((v == VNEC_VAR_S && !slices) || (v == VNEC_VAR_G && !pattern) || (v == VNEC_VAR_D && !dblocks))
            continue;
        double ms = INFINITY;
        rc = time_tuning(a, &FP32_SHAPES[c], dx, dy, 20, &ms);
        if (ms < best) {
            best = ms;
            bi = c;
        }
    }
    free(hx);
    vnec_free_dev(dx);
    vnec_free_dev(dy);
    if (rc != SpB_SUCCESS) return rc;
    a->tuned_fp32 = bi;
    *t = FP32_SHAPES[bi];
    return SpB_SUCCESS;
}

SpB_Info vnec_refine_slice_plan(spb_csr *a, vnec_tuning *t)
{
    const int ipt = t->variant == VNEC_VAR_S ? t->ipt : 3;
    if (a->refined_slice_ipt == ipt) {
        t->variant = a->refined_slice_variant;
        t->ipt = a->refined_slice_ipt_out;
        t->grid = -1;
        return SpB_SUCCESS;
    }
    uint32_t mask = 0;
    if (vnec_candidate_mask_ext((SpB_Matrix)a, &mask) != SpB_SUCCESS) return SpB_FAILURE;
    int64_t nr = (int64_t)a->base.row, nc = (int64_t)a->base.col;
    double *hx = (double *)malloc(sizeof(double) * (size_t)(nc ? nc : 1));
    double *dx = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(nc ? nc : 1), NULL);
    double *dy = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(nr ? nr : 1), NULL);
    SpB_Info rc = (hx && dx && dy) ? SpB_SUCCESS : vnec_fail("out of memory");
    for (int64_t i = 0; rc == SpB_SUCCESS && i < nc; i++) hx[i] = 1.0 / (double)(1 + (i % 7));
    if (rc == SpB_SUCCESS && nc)
        rc = vnec_hip_ok(hipMemcpy(dx, hx, sizeof(double) * (size_t)nc, hipMemcpyHostToDevice), "h2d");
    vnec_tuning s = *t, g = CANDIDATES[CAND_G];
    s.variant = VNEC_VAR_S;
    s.ipt = ipt;
    s.grid = -1;
    g.encoding = t->encoding;
    double ts = INFINITY, tg = INFINITY;
    if (rc == SpB_SUCCESS && (mask & 7u)) rc = time_tuning(a, &s, dx, dy, 20, &ts); /* slices apply */
    if (rc == SpB_SUCCESS && ((mask >> CAND_G) & 1u)) rc = time_tuning(a, &g, dx, dy, 20, &tg);
    free(hx);
    vnec_free_dev(dx);
    vnec_free_dev(dy);
    if (rc != SpB_SUCCESS) return rc;
    if (tg < ts) {
        t->variant = VNEC_VAR_G;
        t->ipt = g.ipt;
    } else if (ts < INFINITY) {
        t->variant = VNEC_VAR_S;
        t->ipt = ipt;
    }
    t->grid = -1;
    a->refined_slice_ipt = ipt;
    a->refined_slice_variant = t->variant;
    a->refined_slice_ipt_out = t->ipt;
    return SpB_SUCCESS;
}

/* Relabelled columns (descending use) are measured as further merge-path
 * forms when x takes at least this many bytes: below it x stays in the
 * caches closest to the kernels whatever its order. */
#define RELABEL_MIN_X_BYTES ((int64_t)1 << 21)

/* Granularity of the 512-thread column-sorted forms measured beside a pick:
 * the pick's own where those tiles exist at it, 9 above it (at 11 items per
 * thread a 512-thread workgroup holds 156 registers per lane, 3 wavefronts
 * per SIMD), 0 where none applies. */
static int wide_ipt(int ipt)
{
    if (ipt > 9) ipt = 9;
    return vnec_hip_stream_supported(VNEC_VAR_C, ipt, VNEC_ENC_FP64, VNEC_WIDE_BLOCK) ? ipt : 0;
}

SpB_Info vnec_refine_merge_plan(spb_csr *a, vnec_tuning *t)
{
    if (a->refined_ipt == t->ipt) {
        t->variant = a->refined_variant;
        t->ipt = a->refined_ipt_out;
        t->grid = a->refined_grid;
        t->lanes = a->refined_lanes;
        return SpB_SUCCESS;
    }
    int64_t nr = (int64_t)a->base.row, nc = (int64_t)a->base.col;
    double *hx = (double *)malloc(sizeof(double) * (size_t)(nc ? nc : 1));
    double *dx = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(nc ? nc : 1), NULL);
    double *dy = (double *)vnec_malloc_dev(sizeof(double) * (size_t)(nr ? nr : 1), NULL);
    SpB_Info rc = (hx && dx && dy) ? SpB_SUCCESS : vnec_fail("out of memory");
    for (int64_t i = 0; rc == SpB_SUCCESS && i < nc; i++) hx[i] = 1.0 / (double)(1 + (i % 7));
    if (rc == SpB_SUCCESS && nc)
        rc = vnec_hip_ok(hipMemcpy(dx, hx, sizeof(double) * (size_t)nc, hipMemcpyHostToDevice), "h2d");
    /* L or C tiles, persistent or with per-tile carries, at the pick's
     * granularity; the same over hot-first relabelled columns when x is
     * large; column-sorted 512-thread tiles, persistent or with per-tile
     * carries, and over relabelled columns with per-tile carries; and those
     * tiles with their gathers issued together, at 7 items per thread or
     * fewer (more registers per thread above) */
    const int relabel = 8 * nc >= RELABEL_MIN_X_BYTES, wipt = wide_ipt(t->ipt);
    const int gipt = t->ipt < 7 ? t->ipt : 7;
    vnec_tuning forms[12];
    int nf = 0;
    for (int k = 0; k < (relabel ? 8 : 4); k++) {
        vnec_tuning u = *t;
        u.variant = (k / 2) % 2 ? VNEC_VAR_C : VNEC_VAR_L;
        u.grid = k % 2 ? -1 : 0;
        u.lanes = k >= 4 ? VNEC_LANES_RELABEL : 64;
        forms[nf++] = u;
    }
    for (int k = 0; wipt && k < (relabel ? 3 : 2); k++) {
        vnec_tuning u = *t;
        u.variant = VNEC_VAR_C;
        u.ipt = wipt;
        u.grid = k == 0 ? 0 : -1;
        u.lanes = k == 2 ? (VNEC_LANES_WIDE | VNEC_LANES_RELABEL) : VNEC_LANES_WIDE;
        forms[nf++] = u;
    }
    if (vnec_hip_strided_supported(VNEC_VAR_C, gipt, VNEC_ENC_FP64, VNEC_WIDE_BLOCK, 1)) {
        vnec_tuning u = *t;
        u.variant = VNEC_VAR_C;
        u.ipt = gipt;
        u.grid = -1;
        u.lanes = VNEC_LANES_WIDE | VNEC_LANES_GATHER;
        forms[nf++] = u;
    }
    double best = INFINITY;
    vnec_tuning b = *t;
    b.lanes = 64;
    for (int k = 0; rc == SpB_SUCCESS && k < nf; k++) {
        double tb = INFINITY;
        rc = time_tuning(a, &forms[k], dx, dy, 20, &tb);
        if (tb < best) {
            best = tb;
            b = forms[k];
        }
    }
    free(hx);
    vnec_free_dev(dx);
    vnec_free_dev(dy);
    if (rc != SpB_SUCCESS) return rc;
    a->refined_ipt = t->ipt;
    a->refined_variant = t->variant = b.variant;
    a->refined_ipt_out = t->ipt = b.ipt;
    a->refined_grid = t->grid = b.grid;
    a->refined_lanes = t->lanes = b.lanes;
    return SpB_SUCCESS;
}

SpB_Info vnec_select_tuning_no_pattern(spb_csr *a, vnec_tuning *t)
{
    int c = CAND_L7_PERSISTENT;
    if (model_trained() && select_learned(a, ~(1u << CAND_G), &c) != SpB_SUCCESS) return SpB_FAILURE;
    if (!model_trained() && a->nnz > 0 && vnec_slice_padding(a) <= RULE_MAX_PADDING) c = CAND_S3;
    *t = CANDIDATES[c];
    return SpB_SUCCESS;
}

SpB_Info vnec_select_tuning(spb_csr *a, int mode, int nrhs, vnec_tuning *t)
{
    if (nrhs > 1) {
        int ipt = 5;
        if (mode == VNEC_SELECT_AUTOTUNE && autotune_multi(a, nrhs, &ipt) != SpB_SUCCESS) return SpB_FAILURE;
        vnec_tuning m = {VNEC_VAR_L, ipt, 64, VNEC_ENC_FP64, nrhs, -1};
        *t = m;
        return SpB_SUCCESS;
    }
    int c = 0;
    if (vnec_select_ext((SpB_Matrix)a, mode, &c) != SpB_SUCCESS) return SpB_FAILURE;
    *t = CANDIDATES[c];
    return SpB_SUCCESS;
}
