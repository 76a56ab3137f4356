// vnec_spmv_d.hip — VNEC-D SpMV for gfx950 (variant D), in FP64 or, for
// the FP32 API, in FP32 (values, x, sums).
//
// The D format (spmv_vnec_D_for_mxv.c:356-392) cuts each row segment of a
// partition into groups of L consecutive positions of the partition's
// compressed x window (its ECR), storing every group as L zero-padded
// values plus one window position.  Here a partition is a merge-path tile
// of 256 * IPT items, so its window fits in LDS, and the value array is
// re-laid for 64-lane wavefronts: each block of 64 consecutive groups is
// stored [j][lane], so lane l of a wave reads value j of group 64b + l and
// the wave's load of value j is one contiguous 512-byte line run.
//
// Per tile (spmv_vnec_D_for_mxv.c:400-485, restated):
//   1. the window x[use_x_indices] is gathered into LDS (coalesced runs),
//      followed by L zeros so that a group reaching past the window's end
//      reads zeros (the reference reads past NEC there);
//   2. each group's dot product over its L window positions is staged in
//      LDS (one group per lane, IPT rounds);
//   3. a merge walk over (row ends in group space, groups) gives every
//      thread IPT items; rows that span threads are joined by the DPP
//      segmented scan; every row x0 <= i < x1 completes in the tile;
//   4. the head of row x1 that lies in the tile, nonzeros
//      [max(ptr[x1], y0), y1), is summed from the CSR values with a
//      fixed-order block reduction and left in carry[t] (the reference's
//      carry, spmv_vnec_D_for_mxv.c:470-477, without its double count
//      when a tile lies inside one row);
//   5. k_carry_fixup adds the carries to y in tile order.
#include <hip/hip_runtime.h>

#include "vnec_common.h"
#include "vnec_hip.h"

template <typename T>
__global__ void k_d_interleave(const double *__restrict__ in, int64_t G, int L, T *__restrict__ out)
{
    const int64_t n = G * L;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i / L, j = i - g * L;
        out[((g >> 6) * L + j) * 64 + (g & 63)] = (T)in[i];
    }
}

// Operands in the kernel's value type (FP64, or FP32 for the FP32 API).
__device__ __forceinline__ const double *dvals_of(const vnec_d_args &a, double) { return a.vals; }
__device__ __forceinline__ const float *dvals_of(const vnec_d_args &a, float) { return a.vals32; }
__device__ __forceinline__ const double *dval_of(const vnec_d_args &a, double) { return a.val; }
__device__ __forceinline__ const float *dval_of(const vnec_d_args &a, float) { return a.val32; }
__device__ __forceinline__ const double *dx_of(const vnec_d_args &a, double) { return a.x; }
__device__ __forceinline__ const float *dx_of(const vnec_d_args &a, float) { return a.x32; }
__device__ __forceinline__ double *dy_of(const vnec_d_args &a, double) { return a.y; }
__device__ __forceinline__ float *dy_of(const vnec_d_args &a, float) { return a.y32; }
__device__ __forceinline__ double *dcarry_of(const vnec_d_args &a, double) { return a.carry; }
__device__ __forceinline__ float *dcarry_of(const vnec_d_args &a, float) { return a.carry32; }

template <int IPT, int L, typename F>
__global__ __launch_bounds__(VNEC_BLOCK) void k_spmv_d(vnec_d_args a, int64_t grid)
{
    constexpr int T = VNEC_BLOCK * IPT;
    __shared__ F s_win[T + L];
    __shared__ F s_prod[T];
    __shared__ int s_rowend[T];
    __shared__ int s_f[VNEC_NWAVES + 1];
    __shared__ F s_v[VNEC_NWAVES + 1];

    const int64_t t = xcd_tile(blockIdx.x, grid);
    if (t >= a.ntiles) return;
    const F *__restrict__ x = dx_of(a, F());
    F *__restrict__ y = dy_of(a, F());
    const int2 *co = (const int2 *)a.coords;
    const int2 c0 = co[t], c1 = co[t + 1];
    const int x0 = c0.x, y0 = c0.y, x1 = c1.x, y1 = c1.y;
    const int nrow = x1 - x0;
    const int tid = threadIdx.x;
    const int nec = a.nec[t];

    // 1. window + L zeros
    fill_window(x, a.run_start, a.run_wpos, a.run_off[t], a.run_off[t + 1], nec, s_win,
                (int *)s_prod);
    if (tid < L) s_win[nec + tid] = 0;
    __syncthreads();

    // 2. group products and row ends (group space)
    const int G0 = a.vrp[x0];
    const int ng = a.vrp[x1] - G0;
    for (int i = tid; i < nrow; i += VNEC_BLOCK) s_rowend[i] = a.vrp[x0 + 1 + i];
#pragma unroll
    for (int u = 0; u < IPT; u++) {
        const int k = u * VNEC_BLOCK + tid;
        if (k < ng) {
            const int64_t g = (int64_t)G0 + k;
            const F *v = dvals_of(a, F()) + ((g >> 6) * L) * 64 + (g & 63);
            const int cs = __builtin_nontemporal_load(a.col_start + g);
            F vv[L];
#pragma unroll
            for (int j = 0; j < L; j++) vv[j] = __builtin_nontemporal_load(v + j * 64);
            F s = 0;
#pragma unroll
            for (int j = 0; j < L; j++) s = fma(vv[j], s_win[cs + j], s);
            s_prod[k] = s;
        }
    }
    __syncthreads();

    // 3. merge walk over (rows, groups)
    const int items = nrow + ng;
    const int d0 = tid * IPT < items ? tid * IPT : items;
    const int d1 = d0 + IPT < items ? d0 + IPT : items;
    int rx = (int)merge_search(s_rowend, nrow, G0, ng, d0);
    int ny = d0 - rx;
    F acc = 0, head_val = 0;
    int flag = 0, head_row = 0;
#pragma unroll
    for (int u = 0; u < IPT; u++) {
        if (d0 + u < d1) {
            if (rx < nrow && s_rowend[rx] <= G0 + ny) {
                if (!flag) {
                    head_row = x0 + rx;
                    head_val = acc;
                    flag = 1;
                } else {
                    y[x0 + rx] = acc;
                }
                acc = 0;
                rx++;
            } else {
                acc += s_prod[ny];
                ny++;
            }
        }
    }
    const F cin = block_seg_carry(flag, acc, s_f, s_v, (F *)nullptr);
    if (flag) y[head_row] = head_val + cin;

    // 4. head of row x1 inside the tile, from the CSR values
    F c = 0;
    if (x1 < a.nrows) {
        const F *__restrict__ val = dval_of(a, F());
        const int b = a.ptr[x1] > y0 ? a.ptr[x1] : y0;
        for (int j = b + tid; j < y1; j += VNEC_BLOCK) c = fma(val[j], s_win[a.lcol[j]], c);
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, VNEC_WAVE);
    __syncthreads();  // s_v is free again
    if (lane_id() == 0) s_v[wave_id()] = c;
    __syncthreads();
    if (tid == 0) {
        F s = 0;
        for (int w = 0; w < VNEC_NWAVES; w++) s += s_v[w];
        dcarry_of(a, F())[t] = s;
    }
}

typedef void (*d_kernel_t)(vnec_d_args, int64_t);

static d_kernel_t pick_d(int ipt, int lanes, bool fp32)
{
#define CASE_D(I, LL) \
    if (ipt == I && lanes == LL) return fp32 ? k_spmv_d<I, LL, float> : k_spmv_d<I, LL, double>;
    CASE_D(5, 4) CASE_D(5, 8) CASE_D(5, 16)
    CASE_D(7, 4) CASE_D(7, 8) CASE_D(7, 16)
    CASE_D(9, 4) CASE_D(9, 8) CASE_D(9, 16)
#undef CASE_D
    return nullptr;
}

extern "C" int vnec_hip_d_supported(int ipt, int lanes) { return pick_d(ipt, lanes, false) != nullptr; }

extern "C" int vnec_hip_d_interleave(const double *in, int64_t G, int L, double *out, void *stream)
{
    if (G <= 0) return 0;
    int64_t n = G * L, blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_d_interleave<double>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       in, G, L, out);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_d_interleave_f32(const double *in, int64_t G, int L, float *out, void *stream)
{
    if (G <= 0) return 0;
    int64_t n = G * L, blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_d_interleave<float>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       in, G, L, out);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_spmv_d(const vnec_d_args *a, void *stream)
{
    d_kernel_t k = pick_d(a->ipt, a->lanes, a->fp32 != 0);
    if (!k) return (int)hipErrorInvalidValue;
    if (a->ntiles <= 0) return 0;
    int64_t grid = xcd_grid(a->ntiles);
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(VNEC_BLOCK), 0, (hipStream_t)stream, *a, grid);
    return (int)hipGetLastError();
}
