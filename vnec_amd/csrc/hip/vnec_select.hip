// vnec_select.hip — kernels of the variant selector (gfx950, wave64):
// integer matrix statistics, and the FP32 MLP's forward, backward, Adam
// update and inference.
//
// Statistics are integer sums, maxima and histograms accumulated with
// integer atomics, so their values do not depend on the order in which
// waves run: a CPU computation of the same definitions matches bit for bit.
//
// The MLP (32 -> 64 -> 64 -> 14, ReLU) is mapped onto one wavefront per
// sample: lane j owns hidden unit j of both hidden layers, and lane c < 14
// owns output c.  Weights are stored input-major (W[k][j]), so for each
// input k the wave reads one contiguous row of 64 floats; activations of
// other lanes are broadcast with readlane.  Gradients are accumulated in
// registers per workgroup over a chunk of samples and reduced over chunks
// in a fixed order, so training is deterministic.
#include "vnec_common.h"
#include "vnec_ext.h"
#include "vnec_hip.h"

#define NF VNEC_NFEAT
#define NH VNEC_MLP_HIDDEN
#define NC VNEC_NCAND
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + NF * NH)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NH * NC)
#define NPARAM (OFF_B3 + NC)
static_assert(NPARAM == VNEC_MLP_NPARAMS, "parameter layout");
static_assert(NH == VNEC_WAVE, "one lane per hidden unit");
static_assert(NC <= VNEC_WAVE && NF <= VNEC_WAVE, "outputs and inputs fit one wave");

#define LEN_BINS 65536

// ------------------------------------------------------------ statistics
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)v, o, VNEC_WAVE);
        uint32_t hi = __shfl_xor((uint32_t)(v >> 32), o, VNEC_WAVE);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v)
{
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)v, o, VNEC_WAVE);
        uint32_t hi = __shfl_xor((uint32_t)(v >> 32), o, VNEC_WAVE);
        uint64_t w = ((uint64_t)hi << 32) | lo;
        v = w > v ? w : v;
    }
    return v;
}

// Row-length statistics.  Rows are visited in aligned groups of 256 (a
// grid-stride loop over blocks), so wave w of a group covers exactly the
// 64-row slice starting at a multiple of 64.
__global__ __launch_bounds__(VNEC_BLOCK) void k_feature_rows(const int32_t *__restrict__ ptr,
                                                             int64_t nrows,
                                                             unsigned long long *__restrict__ st,
                                                             uint32_t *__restrict__ lenhist)
{
    __shared__ uint32_t s_hist[VNEC_FEAT_NBUCKETS];
    if (threadIdx.x < VNEC_FEAT_NBUCKETS) s_hist[threadIdx.x] = 0;
    __syncthreads();
    uint64_t len2 = 0, len3 = 0, mx = 0, empty = 0, slots = 0;
    const int lane = lane_id();
    for (int64_t r0 = (int64_t)blockIdx.x * VNEC_BLOCK; r0 < nrows;
         r0 += (int64_t)gridDim.x * VNEC_BLOCK) {
        const int64_t r = r0 + threadIdx.x;
        const bool live = r < nrows;
        const uint64_t len = live ? (uint64_t)(int64_t)(ptr[r + 1] - ptr[r]) : 0;
        if (live) {
            len2 += len * len;
            len3 += len * len * len;
            mx = len > mx ? len : mx;
            if (len == 0) {
                empty++;
            } else {
                int b = 63 - __clzll((long long)len);
                atomicAdd(&s_hist[b < VNEC_FEAT_NBUCKETS - 1 ? b : VNEC_FEAT_NBUCKETS - 1], 1u);
            }
        }
        const uint64_t sm = wave_max_u64(len);
        if (lane == 0) slots += 64 * sm;
        // exact histogram of min(len, 65535), one atomic per distinct
        // length present in the wave
        const uint32_t key = len < LEN_BINS - 1 ? (uint32_t)len : LEN_BINS - 1;
        bool pending = live;
        for (;;) {
            const uint64_t am = __ballot(pending);
            if (am == 0) break;
            const int leader = __ffsll((unsigned long long)am) - 1;
            const uint32_t lk = (uint32_t)__shfl((int)key, leader, VNEC_WAVE);
            const uint64_t same = __ballot(pending && key == lk);
            if (lane == leader) atomicAdd(&lenhist[lk], (uint32_t)__popcll(same));
            if (key == lk) pending = false;
        }
    }
    len2 = wave_sum_u64(len2);
    len3 = wave_sum_u64(len3);
    empty = wave_sum_u64(empty);
    mx = wave_max_u64(mx);
    if (lane == 0) {
        atomicAdd(&st[VF_SUM_LEN2], (unsigned long long)len2);
        atomicAdd(&st[VF_SUM_LEN3], (unsigned long long)len3);
        atomicAdd(&st[VF_EMPTY_ROWS], (unsigned long long)empty);
        atomicAdd(&st[VF_SLICE_SLOTS], (unsigned long long)slots);
        atomicMax(&st[VF_MAX_LEN], (unsigned long long)mx);
    }
    __syncthreads();
    if (threadIdx.x < VNEC_FEAT_NBUCKETS && s_hist[threadIdx.x])
        atomicAdd(&st[VF_HIST + threadIdx.x], (unsigned long long)s_hist[threadIdx.x]);
}

// Distance of every entry from the scaled diagonal row * ncols / nrows:
// one wave per row (grid-stride), lanes over the row's entries.
__global__ __launch_bounds__(VNEC_BLOCK) void k_feature_cols(const int32_t *__restrict__ ptr,
                                                             const int32_t *__restrict__ col,
                                                             int64_t nrows, int64_t ncols,
                                                             unsigned long long *__restrict__ st)
{
    uint64_t sum = 0, mx = 0, near = 0;
    const int lane = lane_id();
    const int64_t nw = (int64_t)gridDim.x * VNEC_NWAVES;
    for (int64_t r = (int64_t)blockIdx.x * VNEC_NWAVES + wave_id(); r < nrows; r += nw) {
        const int64_t diag = r * ncols / nrows;
        const int32_t b = ptr[r], e = ptr[r + 1];
        for (int32_t j = b + lane; j < e; j += VNEC_WAVE) {
            int64_t d = (int64_t)col[j] - diag;
            const uint64_t ad = (uint64_t)(d < 0 ? -d : d);
            sum += ad;
            mx = ad > mx ? ad : mx;
            near += ad <= 32;
        }
    }
    sum = wave_sum_u64(sum);
    near = wave_sum_u64(near);
    mx = wave_max_u64(mx);
    if (lane == 0) {
        atomicAdd(&st[VF_SUM_DIST], (unsigned long long)sum);
        atomicAdd(&st[VF_NEAR_DIAG], (unsigned long long)near);
        atomicMax(&st[VF_MAX_DIST], (unsigned long long)mx);
    }
}

extern "C" int vnec_hip_feature_rows(const int32_t *ptr, int64_t nrows, uint64_t *stats,
                                     uint32_t *lenhist, void *stream)
{
    if (nrows <= 0) return 0;
    int64_t blocks = (nrows + VNEC_BLOCK - 1) / VNEC_BLOCK;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_feature_rows, dim3((unsigned)blocks), dim3(VNEC_BLOCK), 0,
                       (hipStream_t)stream, ptr, nrows, (unsigned long long *)stats, lenhist);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_feature_cols(const int32_t *ptr, const int32_t *col, int64_t nrows,
                                     int64_t ncols, uint64_t *stats, void *stream)
{
    if (nrows <= 0) return 0;
    int64_t blocks = (nrows + VNEC_NWAVES - 1) / VNEC_NWAVES;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_feature_cols, dim3((unsigned)blocks), dim3(VNEC_BLOCK), 0,
                       (hipStream_t)stream, ptr, col, nrows, ncols, (unsigned long long *)stats);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------------- MLP
__device__ __forceinline__ float lane_bcast(float v, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

struct MlpAct {
    float a1, h1, a2, h2, o; // pre-activations, activations, output (lane < NC)
};

// Forward pass of one sample whose (standardised) inputs are x[0 .. NF).
__device__ __forceinline__ MlpAct mlp_forward(const float *__restrict__ P, const float *x)
{
    const int j = lane_id();
    MlpAct f;
    float a = P[OFF_B1 + j];
#pragma unroll
    for (int k = 0; k < NF; k++) a = fmaf(P[OFF_W1 + k * NH + j], x[k], a);
    f.a1 = a;
    f.h1 = a > 0.f ? a : 0.f;
    a = P[OFF_B2 + j];
#pragma unroll
    for (int k = 0; k < NH; k++) a = fmaf(P[OFF_W2 + k * NH + j], lane_bcast(f.h1, k), a);
    f.a2 = a;
    f.h2 = a > 0.f ? a : 0.f;
    const int c = j < NC ? j : NC - 1;
    a = P[OFF_B3 + c];
#pragma unroll
    for (int k = 0; k < NH; k++) a = fmaf(P[OFF_W3 + k * NC + c], lane_bcast(f.h2, k), a);
    f.o = a;
    return f;
}

__global__ __launch_bounds__(VNEC_WAVE) void k_mlp_forward(const float *__restrict__ P,
                                                           const float *__restrict__ X, int64_t B,
                                                           float *__restrict__ out)
{
    const int64_t b = blockIdx.x;
    if (b >= B) return;
    MlpAct f = mlp_forward(P, X + b * NF);
    if (lane_id() < NC) out[b * NC + lane_id()] = f.o;
}

// Gradient of sum_{s, c} mask (o - T)^2 * inv_den over one chunk of
// samples per workgroup (one wavefront).
__global__ __launch_bounds__(VNEC_WAVE) void k_mlp_grad(
    const float *__restrict__ P, const float *__restrict__ X, const float *__restrict__ T,
    const float *__restrict__ M, const int32_t *__restrict__ idx, int64_t B, int chunk,
    float inv_den, float *__restrict__ part, float *__restrict__ loss_part)
{
    const int j = lane_id();
    float gW1[NF], gW2[NH], gW3[NH];
#pragma unroll
    for (int k = 0; k < NF; k++) gW1[k] = 0.f;
#pragma unroll
    for (int k = 0; k < NH; k++) {
        gW2[k] = 0.f;
        gW3[k] = 0.f;
    }
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, lsum = 0.f;
    const int64_t s0 = (int64_t)blockIdx.x * chunk;
    const int64_t s1 = s0 + chunk < B ? s0 + chunk : B;
    for (int64_t s = s0; s < s1; s++) {
        const int64_t sid = idx ? (int64_t)idx[s] : s;
        const float *x = X + sid * NF;
        const MlpAct f = mlp_forward(P, x);
        const float t = j < NC ? T[sid * NC + j] : 0.f;
        const float m = j < NC ? M[sid * NC + j] : 0.f;
        const float diff = f.o - t;
        const float dO = j < NC ? 2.f * m * diff * inv_den : 0.f;
        lsum = fmaf(m * diff, diff, lsum);
        gb3 += dO;
#pragma unroll
        for (int k = 0; k < NH; k++) gW3[k] = fmaf(lane_bcast(f.h2, k), dO, gW3[k]);
        float d2 = 0.f;
#pragma unroll
        for (int c = 0; c < NC; c++) d2 = fmaf(P[OFF_W3 + j * NC + c], lane_bcast(dO, c), d2);
        d2 = f.a2 > 0.f ? d2 : 0.f;
        gb2 += d2;
#pragma unroll
        for (int k = 0; k < NH; k++) gW2[k] = fmaf(lane_bcast(f.h1, k), d2, gW2[k]);
        float d1 = 0.f;
#pragma unroll
        for (int k = 0; k < NH; k++) d1 = fmaf(P[OFF_W2 + j * NH + k], lane_bcast(d2, k), d1);
        d1 = f.a1 > 0.f ? d1 : 0.f;
        gb1 += d1;
#pragma unroll
        for (int k = 0; k < NF; k++) gW1[k] = fmaf(x[k], d1, gW1[k]);
    }
    float *out = part + (int64_t)blockIdx.x * NPARAM;
#pragma unroll
    for (int k = 0; k < NF; k++) out[OFF_W1 + k * NH + j] = gW1[k];
    out[OFF_B1 + j] = gb1;
#pragma unroll
    for (int k = 0; k < NH; k++) out[OFF_W2 + k * NH + j] = gW2[k];
    out[OFF_B2 + j] = gb2;
    if (j < NC) {
#pragma unroll
        for (int k = 0; k < NH; k++) out[OFF_W3 + k * NC + j] = gW3[k];
        out[OFF_B3 + j] = gb3;
    }
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, VNEC_WAVE);
    if (j == 0) loss_part[blockIdx.x] = lsum * inv_den;
}

__global__ __launch_bounds__(VNEC_BLOCK) void k_mlp_reduce_adam(
    const float *__restrict__ part, int nchunks, float *__restrict__ grad, float *__restrict__ P,
    float *__restrict__ m, float *__restrict__ v, float lr, float wd, float bc1, float bc2)
{
    const int p = blockIdx.x * VNEC_BLOCK + threadIdx.x;
    if (p >= NPARAM) return;
    float g = 0.f;
    for (int c = 0; c < nchunks; c++) g += part[(int64_t)c * NPARAM + p];
    if (grad) grad[p] = g;
    if (lr > 0.f) {
        const float w = P[p];
        g = fmaf(wd, w, g);
        const float mm = 0.9f * m[p] + 0.1f * g;
        const float vv = 0.999f * v[p] + 0.001f * g * g;
        m[p] = mm;
        v[p] = vv;
        P[p] = w - lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
    }
}

__global__ __launch_bounds__(VNEC_WAVE) void k_mlp_select(
    const float *__restrict__ P, const float *__restrict__ mu, const float *__restrict__ sd,
    const float *__restrict__ F, const uint32_t *__restrict__ mask, int64_t B,
    int32_t *__restrict__ choice)
{
    __shared__ float s_x[NF];
    const int64_t b = blockIdx.x;
    if (b >= B) return;
    const int j = lane_id();
    if (j < NF) s_x[j] = (F[b * NF + j] - mu[j]) / sd[j];
    __syncthreads();
    const MlpAct f = mlp_forward(P, s_x);
    const uint32_t mk = mask[b];
    const bool ok = j < NC && ((mk >> j) & 1u) && f.o == f.o;
    float val = ok ? f.o : INFINITY;
    int id = ok ? j : 0x7fffffff;
    for (int o = 32; o > 0; o >>= 1) {
        const float v2 = __shfl_xor(val, o, VNEC_WAVE);
        const int i2 = __shfl_xor(id, o, VNEC_WAVE);
        if (v2 < val || (v2 == val && i2 < id)) {
            val = v2;
            id = i2;
        }
    }
    if (j == 0) choice[b] = id == 0x7fffffff ? -1 : id;
}

extern "C" int vnec_hip_mlp_forward(const float *params, const float *X, int64_t B, float *out,
                                    void *stream)
{
    if (B <= 0) return 0;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)B), dim3(VNEC_WAVE), 0, (hipStream_t)stream,
                       params, X, B, out);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_mlp_grad(const float *params, const float *X, const float *T,
                                 const float *mask, const int32_t *idx, int64_t B, int chunk,
                                 float inv_den, float *part, float *loss_part, void *stream)
{
    if (B <= 0 || chunk <= 0) return (int)hipErrorInvalidValue;
    int64_t nch = (B + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)nch), dim3(VNEC_WAVE), 0, (hipStream_t)stream,
                       params, X, T, mask, idx, B, chunk, inv_den, part, loss_part);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_mlp_reduce_adam(const float *part, int nchunks, float *grad, float *params,
                                        float *m, float *v, float lr, float weight_decay,
                                        float bc1, float bc2, void *stream)
{
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((NPARAM + VNEC_BLOCK - 1) / VNEC_BLOCK),
                       dim3(VNEC_BLOCK), 0, (hipStream_t)stream, part, nchunks, grad, params, m, v,
                       lr, weight_decay, bc1, bc2);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_mlp_select(const float *params, const float *mu, const float *sd,
                                   const float *features, const uint32_t *mask, int64_t B,
                                   int32_t *choice, void *stream)
{
    if (B <= 0) return 0;
    hipLaunchKernelGGL(k_mlp_select, dim3((unsigned)B), dim3(VNEC_WAVE), 0, (hipStream_t)stream,
                       params, mu, sd, features, mask, B, choice);
    return (int)hipGetLastError();
}
