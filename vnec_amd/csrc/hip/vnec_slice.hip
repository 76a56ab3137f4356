// vnec_slice.hip — VNEC-S re-laid for 64-lane wavefronts (variant S).
//
// The reference's S format (spmv_vnec_S_for_mxv.c:442-585) groups L = 8
// rows of equal length and stores each group transposed, [l][row], so one
// SIMD load serves eight rows; rows that do not fill a group are handled
// one at a time.  On gfx950 the group is a wavefront: a slice holds 64
// consecutive rows, padded to the slice's longest row, stored [l][lane].
// Lane `lane` of a wavefront owns row 64 s + lane and accumulates it in a
// register: every value / column load is one coalesced 512 B / 128 B
// access, and no cross-lane reduction is needed.  Padding slots are never
// read (each lane stops at its own row length).
//
// Values are FP64, FP32 hi + lo or FP32 hi (summed in FP64), or, for the
// FP32 API, FP32 with FP32 x and sums.
//
// Tiles are runs of whole slices, balanced by the merge-path partitioner
// over the slice layout (row snap on).  Each tile's x window (its distinct
// columns, ECR) is gathered into LDS and columns are stored as 16-bit
// window ranks; when a tile's window is too large the plan stores global
// int32 columns instead and gathers x directly.
#include <hip/hip_runtime.h>

#include "vnec_common.h"
#include "vnec_hip.h"

#define SLICE 64

// ------------------------------------------------------------- builders
// slice_len[s] = longest row of slice s; slice_ptr gets 64 * slice_len
// (scanned by the caller into the padded slice offsets).
__global__ void k_slice_len(const int32_t *__restrict__ ptr, int64_t nrows, int64_t nslices,
                            int32_t *__restrict__ slice_len, int32_t *__restrict__ slice_sz)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s <= nslices;
         s += (int64_t)gridDim.x * blockDim.x) {
        if (s == nslices) {
            slice_sz[s] = 0;
            continue;
        }
        int64_t r0 = s * SLICE, r1 = r0 + SLICE < nrows ? r0 + SLICE : nrows;
        int m = 0;
        for (int64_t r = r0; r < r1; r++) {
            int l = ptr[r + 1] - ptr[r];
            m = l > m ? l : m;
        }
        slice_len[s] = m;
        slice_sz[s] = m * SLICE;
    }
}

// Tile boundaries in (row, nonzero) space from slice-space coordinates.
__global__ void k_slice_tile_rows(const int2 *__restrict__ sco, int64_t ntiles,
                                  const int32_t *__restrict__ ptr, int64_t nrows,
                                  int2 *__restrict__ rco)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > ntiles) return;
    int64_t r = (int64_t)sco[t].x * SLICE;
    if (r > nrows) r = nrows;
    rco[t] = make_int2((int)r, ptr[r]);
}

// One wavefront per slice: lane = row.  Writes the transposed, padded
// value (FP64 or FP32 hi / lo) and column (global int32 or 16-bit window
// rank) arrays.  Padding slots get value 0 and column 0.
__global__ __launch_bounds__(VNEC_BLOCK) void k_slice_fill(
    const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
    const uint16_t *__restrict__ lcol, const double *__restrict__ val, int64_t nrows,
    int64_t nslices, const int32_t *__restrict__ slice_len, const int32_t *__restrict__ slice_ptr,
    double *__restrict__ sv, float *__restrict__ shi, float *__restrict__ slo,
    int32_t *__restrict__ scol, uint16_t *__restrict__ slcol)
{
    int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / VNEC_WAVE;
    if (s >= nslices) return;
    int lane = lane_id();
    int64_t row = s * SLICE + lane;
    int len = slice_len[s];
    int64_t base = slice_ptr[s];
    int64_t b = 0, rl = 0;
    if (row < nrows) {
        b = ptr[row];
        rl = ptr[row + 1] - b;
    }
    for (int l = 0; l < len; l++) {
        int64_t slot = base + (int64_t)l * SLICE + lane;
        double a = 0.0;
        int c = 0;
        if (l < rl) {
            a = val[b + l];
            c = lcol ? (int)lcol[b + l] : col[b + l];
        }
        if (sv) sv[slot] = a;
        if (shi) {
            float h = (float)a;
            shi[slot] = h;
            if (slo) slo[slot] = (float)(a - (double)h);
        }
        if (scol) scol[slot] = c;
        if (slcol) slcol[slot] = (uint16_t)c;
    }
}

extern "C" int vnec_hip_slice_len(const int32_t *ptr, int64_t nrows, int64_t nslices,
                                  int32_t *slice_len, int32_t *slice_sz, void *stream)
{
    int64_t n = nslices + 1;
    unsigned g = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    hipLaunchKernelGGL(k_slice_len, dim3(g), dim3(256), 0, (hipStream_t)stream, ptr, nrows, nslices,
                       slice_len, slice_sz);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_slice_tile_rows(const int32_t *slice_coords, int64_t ntiles,
                                        const int32_t *ptr, int64_t nrows, int32_t *row_coords,
                                        void *stream)
{
    unsigned g = (unsigned)((ntiles + 1 + 255) / 256);
    hipLaunchKernelGGL(k_slice_tile_rows, dim3(g), dim3(256), 0, (hipStream_t)stream,
                       (const int2 *)slice_coords, ntiles, ptr, nrows, (int2 *)row_coords);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_slice_fill(const int32_t *ptr, const int32_t *col, const uint16_t *lcol,
                                   const double *val, int64_t nrows, int64_t nslices,
                                   const int32_t *slice_len, const int32_t *slice_ptr, double *sv,
                                   float *shi, float *slo, int32_t *scol, uint16_t *slcol,
                                   void *stream)
{
    if (nslices <= 0) return 0;
    int64_t threads = nslices * VNEC_WAVE;
    unsigned g = (unsigned)((threads + VNEC_BLOCK - 1) / VNEC_BLOCK);
    hipLaunchKernelGGL(k_slice_fill, dim3(g), dim3(VNEC_BLOCK), 0, (hipStream_t)stream, ptr, col,
                       lcol, val, nrows, nslices, slice_len, slice_ptr, sv, shi, slo, scol, slcol);
    return (int)hipGetLastError();
}

// --------------------------------------------------------------- SpMV
template <int ENC>
__device__ __forceinline__ typename enc_real<ENC>::type slice_value(const vnec_slice_args &a, int64_t slot)
{
    if (ENC == 0) return __builtin_nontemporal_load(a.sv + slot);
    if (ENC == 1)
        return (double)__builtin_nontemporal_load(a.shi + slot) +
               (double)__builtin_nontemporal_load(a.slo + slot);
    return __builtin_nontemporal_load(a.shi + slot);  // FP32 hi (FP64 sums), or FP32 values
}

__device__ __forceinline__ const double *sx_of(const vnec_slice_args &a, double) { return a.x; }
__device__ __forceinline__ const float *sx_of(const vnec_slice_args &a, float) { return a.x32; }
__device__ __forceinline__ double *sy_of(const vnec_slice_args &a, double) { return a.y; }
__device__ __forceinline__ float *sy_of(const vnec_slice_args &a, float) { return a.y32; }

// One workgroup per tile; wavefront w takes slices w, w + 4, ... of it.
template <int ENC, bool WIN>
__global__ __launch_bounds__(VNEC_BLOCK) void k_spmv_slice(vnec_slice_args a, int64_t grid)
{
    typedef typename enc_real<ENC>::type F;
    extern __shared__ double s_dyn[];
    F *s_win = (F *)s_dyn;
    const int64_t t = xcd_tile(blockIdx.x, grid);
    if (t >= a.ntiles) return;
    const F *__restrict__ x = sx_of(a, F());
    const int2 *co = (const int2 *)a.tile_slices;
    const int sb = co[t].x, se = co[t + 1].x;
    if (WIN) {
        fill_window(x, a.run_start, a.run_wpos, a.run_off[t], a.run_off[t + 1], a.nec[t], s_win,
                    (int *)nullptr);
        __syncthreads();
    }
    const int lane = lane_id();
    const int w = __builtin_amdgcn_readfirstlane(wave_id());
    for (int s = sb + w; s < se; s += VNEC_NWAVES) {
        const int64_t row = (int64_t)s * SLICE + lane;
        int rl = 0;
        if (row < a.nrows) rl = a.ptr[row + 1] - a.ptr[row];
        const int len = a.slice_len[s];
        const int64_t base = (int64_t)a.slice_ptr[s] + lane;
        F acc = 0;
        int l = 0;
        for (; l + 4 <= len; l += 4) {
            F v[4];
            int c[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (l + u < rl) {
                    const int64_t slot = base + (int64_t)(l + u) * SLICE;
                    v[u] = slice_value<ENC>(a, slot);
                    c[u] = WIN ? (int)__builtin_nontemporal_load(a.slcol + slot)
                               : __builtin_nontemporal_load(a.scol + slot);
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (l + u < rl) acc += v[u] * (WIN ? s_win[c[u]] : x[c[u]]);
        }
        for (; l < len; l++) {
            if (l < rl) {
                const int64_t slot = base + (int64_t)l * SLICE;
                const F v = slice_value<ENC>(a, slot);
                const int c = WIN ? (int)__builtin_nontemporal_load(a.slcol + slot)
                                  : __builtin_nontemporal_load(a.scol + slot);
                acc += v * (WIN ? s_win[c] : x[c]);
            }
        }
        if (row < a.nrows) __builtin_nontemporal_store(acc, sy_of(a, F()) + row);  // not read back here
    }
}

typedef void (*slice_kernel_t)(vnec_slice_args, int64_t);

static slice_kernel_t pick_slice(int enc, bool win)
{
    if (win) {
        if (enc == 0) return k_spmv_slice<0, true>;
        if (enc == 1) return k_spmv_slice<1, true>;
        if (enc == 2) return k_spmv_slice<2, true>;
        if (enc == VNEC_ENC_F32_K) return k_spmv_slice<VNEC_ENC_F32_K, true>;
    } else {
        if (enc == 0) return k_spmv_slice<0, false>;
        if (enc == 1) return k_spmv_slice<1, false>;
        if (enc == 2) return k_spmv_slice<2, false>;
        if (enc == VNEC_ENC_F32_K) return k_spmv_slice<VNEC_ENC_F32_K, false>;
    }
    return nullptr;
}

extern "C" int vnec_hip_spmv_slice(const vnec_slice_args *a, void *stream)
{
    slice_kernel_t k = pick_slice(a->encoding, a->win != 0);
    if (!k) return (int)hipErrorInvalidValue;
    if (a->ntiles <= 0) return 0;
    int64_t grid = xcd_grid(a->ntiles);
    size_t esz = a->encoding == VNEC_ENC_F32_K ? sizeof(float) : sizeof(double);
    size_t lds = a->win ? esz * (size_t)(a->max_nec > 0 ? a->max_nec : 1) : 0;
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(VNEC_BLOCK), lds, (hipStream_t)stream, *a, grid);
    return (int)hipGetLastError();
}
