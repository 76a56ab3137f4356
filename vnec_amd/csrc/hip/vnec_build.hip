// vnec_build.hip — GPU builders: merge-path partitioner, exclusive scan,
// ECR column compression (per-tile LDS sort and dense-flag forms), the
// column-sorted tiles of variant C, the right-hand-side layout transposes,
// the FP32 hi/lo value split and small utilities (fills, widening, the
// gather-ceiling probe).
#include <hip/hip_runtime.h>

#include <climits>

#include "vnec_common.h"
#include "vnec_hip.h"

#define HIPCHK(x)                               \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) return (int)e_;   \
    } while (0)

// ------------------------------------------------------------ partitioner
// One lane per diagonal d_t = min(t * ipt, M), ipt = ceil(M / P): the
// binary search of MergePathDivide (spmv_vnec_common.c:3-33) on the row
// ends ptr[1..nrows] against 0..nnz-1; snap applies the S row alignment
// of spmv_vnec_S_for_mxv.c:363-370 (x + 1 clamped to nrows, y = ptr[x]).
__global__ void k_partition(const int32_t *__restrict__ ptr, int64_t nrows, int64_t nnz, int64_t P,
                            int64_t ipt, int snap, int2 *__restrict__ coords)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t > P) return;
    int64_t M = nrows + nnz;
    int64_t d = t * ipt < M ? t * ipt : M;
    int64_t x = merge_search(ptr + 1, nrows, 0, nnz, d);
    int64_t y = d - x;
    if (snap) {
        if (t == 0)
            x = 0;
        else if (t == P)
            x = nrows;
        else
            x = x + 1 < nrows ? x + 1 : nrows;
        y = ptr[x];
    }
    coords[t] = make_int2((int)x, (int)y);
}

extern "C" int vnec_hip_partition(const int32_t *ptr, int64_t nrows, int64_t nnz, int64_t P,
                                  int snap, int32_t *coords, void *stream)
{
    if (P < 1) return (int)hipErrorInvalidValue;
    int64_t M = nrows + nnz;
    int64_t ipt = (M + P - 1) / P;
    int64_t n = P + 1;
    int64_t grid = (n + 255) / 256;
    hipLaunchKernelGGL(k_partition, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, ptr,
                       nrows, nnz, P, ipt, snap, (int2 *)coords);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------------ scan
#define SCAN_ITEMS 8
#define SCAN_TILE (VNEC_BLOCK * SCAN_ITEMS)

__global__ __launch_bounds__(VNEC_BLOCK) void k_scan_reduce(const int32_t *__restrict__ in, int64_t n,
                                                            int32_t *__restrict__ sums)
{
    __shared__ int s_tmp[VNEC_NWAVES + 1];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int v = 0;
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; u++) {
        int64_t i = base + (int64_t)u * VNEC_BLOCK + threadIdx.x;
        if (i < n) v += in[i];
    }
    int tot;
    block_excl_scan_i32(v, s_tmp, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Exclusive scan of one tile (thread-contiguous items), plus offset.
__global__ __launch_bounds__(VNEC_BLOCK) void k_scan_apply(const int32_t *__restrict__ in, int32_t *out,
                                                           int64_t n, const int32_t *__restrict__ offs)
{
    __shared__ int s_tmp[VNEC_NWAVES + 1];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int run = 0;
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; u++) {
        int64_t i = base + u;
        v[u] = i < n ? in[i] : 0;
        run += v[u];
    }
    int tot;
    int ex = block_excl_scan_i32(run, s_tmp, &tot) + (offs ? offs[blockIdx.x] : 0);
#pragma unroll
    for (int u = 0; u < SCAN_ITEMS; u++) {
        int64_t i = base + u;
        if (i < n) out[i] = ex;
        ex += v[u];
    }
}

static int scan_rec(const int32_t *in, int32_t *out, int64_t n, hipStream_t s)
{
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb <= 1) {
        hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(VNEC_BLOCK), 0, s, in, out, n,
                           (const int32_t *)nullptr);
        return (int)hipGetLastError();
    }
    int32_t *sums = nullptr;
    HIPCHK(hipMalloc(&sums, sizeof(int32_t) * (size_t)nb));
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(VNEC_BLOCK), 0, s, in, n, sums);
    int rc = scan_rec(sums, sums, nb, s);
    if (rc == 0) {
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(VNEC_BLOCK), 0, s, in, out, n,
                           (const int32_t *)sums);
        rc = (int)hipGetLastError();
    }
    hipError_t e = hipStreamSynchronize(s);
    hipError_t f = hipFree(sums);
    return rc ? rc : (e ? (int)e : (int)f);
}

extern "C" int vnec_hip_scan_excl_i32(const int32_t *in, int32_t *out, int64_t n,
                                      int64_t *total_host, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (n <= 0) {
        if (total_host) *total_host = 0;
        return 0;
    }
    int32_t last_in = 0;
    if (total_host) HIPCHK(hipMemcpyAsync(&last_in, in + n - 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int rc = scan_rec(in, out, n, s);
    if (rc) return rc;
    if (total_host) {
        int32_t last_out = 0;
        HIPCHK(hipMemcpyAsync(&last_out, out + n - 1, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        *total_host = (int64_t)last_out + last_in;
    }
    return 0;
}

// ------------------------------------------------------- ECR, tile form
// One workgroup per tile: sort the tile's columns (with their positions)
// in LDS, mark first occurrences, scan -> rank of every nonzero among the
// tile's distinct columns (spmv_vnec_D_for_mxv.c:292-319 computes the same
// ranks with dense flags).  Runs of consecutive distinct columns are
// counted for the compact window description.
template <int CAP>
__global__ __launch_bounds__(VNEC_BLOCK) void k_ecr_tile_pass1(
    const int32_t *__restrict__ col, const int2 *__restrict__ coords, int64_t ntiles,
    uint16_t *__restrict__ lcol16, int32_t *__restrict__ ecr32, int32_t *__restrict__ uniq,
    int32_t *__restrict__ nec, int32_t *__restrict__ nruns)
{
    __shared__ uint32_t s_key[CAP];
    __shared__ uint16_t s_idx[CAP];
    __shared__ int s_tmp[VNEC_NWAVES + 1];
    int64_t t = blockIdx.x;
    if (t >= ntiles) return;
    int y0 = coords[t].y, y1 = coords[t + 1].y;
    int n = y1 - y0;
    int n2 = 1;
    while (n2 < n) n2 <<= 1;
    for (int k = threadIdx.x; k < n2; k += VNEC_BLOCK) {
        s_key[k] = k < n ? (uint32_t)col[y0 + k] : 0xFFFFFFFFu;
        s_idx[k] = (uint16_t)k;
    }
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += VNEC_BLOCK) {
                int ixj = i ^ j;
                if (ixj > i) {
                    bool up = (i & k) == 0;
                    uint32_t a = s_key[i], b = s_key[ixj];
                    if ((a > b) == up) {
                        s_key[i] = b;
                        s_key[ixj] = a;
                        uint16_t ti = s_idx[i];
                        s_idx[i] = s_idx[ixj];
                        s_idx[ixj] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    int uniq_base = 0, run_base = 0;
    for (int base = 0; base < n; base += VNEC_BLOCK) {
        int k = base + threadIdx.x;
        int f = 0, g = 0;
        uint32_t key = 0;
        if (k < n) {
            key = s_key[k];
            f = (k == 0) || (key != s_key[k - 1]);
            g = f && ((k == 0) || (key != s_key[k - 1] + 1u));
        }
        int totf, totg;
        int exf = block_excl_scan_i32(f, s_tmp, &totf);
        block_excl_scan_i32(g, s_tmp, &totg);
        if (k < n) {
            int rank = uniq_base + exf + f - 1;
            int pos = y0 + s_idx[k];
            if (lcol16) lcol16[pos] = (uint16_t)rank;
            if (ecr32) ecr32[pos] = rank;
            if (f) uniq[y0 + rank] = (int32_t)key;
        }
        uniq_base += totf;
        run_base += totg;
    }
    if (threadIdx.x == 0) {
        nec[t] = uniq_base;
        nruns[t] = run_base;
    }
}

__global__ __launch_bounds__(VNEC_BLOCK) void k_ecr_tile_pass2(
    const int2 *__restrict__ coords, int64_t ntiles, const int32_t *__restrict__ uniq,
    const int32_t *__restrict__ nec, const int32_t *__restrict__ run_off,
    int32_t *__restrict__ run_start, int32_t *__restrict__ run_wpos)
{
    __shared__ int s_tmp[VNEC_NWAVES + 1];
    int64_t t = blockIdx.x;
    if (t >= ntiles) return;
    int y0 = coords[t].y;
    int n = nec[t];
    int out = run_off[t];
    for (int base = 0; base < n; base += VNEC_BLOCK) {
        int r = base + threadIdx.x;
        int g = 0, u = 0;
        if (r < n) {
            u = uniq[y0 + r];
            g = (r == 0) || (u != uniq[y0 + r - 1] + 1);
        }
        int tot;
        int ex = block_excl_scan_i32(g, s_tmp, &tot);
        if (g) {
            run_start[out + ex] = u;
            run_wpos[out + ex] = r;
        }
        out += tot;
    }
}

extern "C" int vnec_hip_ecr_tiles_pass1(const int32_t *col, const int32_t *coords, int64_t ntiles,
                                        uint16_t *lcol16, int32_t *ecr32, int32_t *uniq,
                                        int32_t *nec, int32_t *nruns, void *stream)
{
    if (ntiles <= 0) return 0;
    int32_t mx = 0;
    int rc = vnec_hip_max_tile_nnz(coords, ntiles, &mx, stream);
    if (rc) return rc;
    if (mx > VNEC_ECR_TILE_MAX_BIG) return (int)hipErrorInvalidValue;
    if (mx <= VNEC_ECR_TILE_MAX)
        hipLaunchKernelGGL(k_ecr_tile_pass1<VNEC_ECR_TILE_MAX>, dim3((unsigned)ntiles), dim3(VNEC_BLOCK),
                           0, (hipStream_t)stream, col, (const int2 *)coords, ntiles, lcol16, ecr32,
                           uniq, nec, nruns);
    else
        hipLaunchKernelGGL(k_ecr_tile_pass1<VNEC_ECR_TILE_MAX_BIG>, dim3((unsigned)ntiles),
                           dim3(VNEC_BLOCK), 0, (hipStream_t)stream, col, (const int2 *)coords, ntiles,
                           lcol16, ecr32, uniq, nec, nruns);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_ecr_tiles_pass2(const int32_t *coords, int64_t ntiles, const int32_t *uniq,
                                        const int32_t *nec, const int32_t *run_off,
                                        int32_t *run_start, int32_t *run_wpos, void *stream)
{
    if (ntiles <= 0) return 0;
    hipLaunchKernelGGL(k_ecr_tile_pass2, dim3((unsigned)ntiles), dim3(VNEC_BLOCK), 0,
                       (hipStream_t)stream, (const int2 *)coords, ntiles, uniq, nec, run_off,
                       run_start, run_wpos);
    return (int)hipGetLastError();
}

__global__ __launch_bounds__(VNEC_BLOCK) void k_max_tile_nnz(const int2 *__restrict__ coords,
                                                             int64_t ntiles, int32_t *out)
{
    int m = 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ntiles;
         t += (int64_t)gridDim.x * blockDim.x) {
        int n = coords[t + 1].y - coords[t].y;
        m = n > m ? n : m;
    }
    for (int o = 32; o > 0; o >>= 1) {
        int v = __shfl_down(m, o, VNEC_WAVE);
        m = v > m ? v : m;
    }
    if (lane_id() == 0) atomicMax(out, m);
}

extern "C" int vnec_hip_max_tile_nnz(const int32_t *coords, int64_t ntiles, int32_t *max_host,
                                     void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    int32_t *d = nullptr;
    HIPCHK(hipMalloc(&d, 4));
    HIPCHK(hipMemsetAsync(d, 0, 4, s));
    int64_t grid = (ntiles + 255) / 256;
    if (grid > 1024) grid = 1024;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_max_tile_nnz, dim3((unsigned)grid), dim3(VNEC_BLOCK), 0, s,
                       (const int2 *)coords, ntiles, d);
    int rc = (int)hipGetLastError();
    if (!rc) rc = (int)hipMemcpyAsync(max_host, d, 4, hipMemcpyDeviceToHost, s);
    if (!rc) rc = (int)hipStreamSynchronize(s);
    hipError_t f = hipFree(d);
    return rc ? rc : (int)f;
}

// ------------------------------------------------------ ECR, dense form
// The reference's own formulation (spmv_vnec_D_for_mxv.c:283-329): a flag
// per (partition, column), a prefix count, and the rank read back per
// nonzero.  One global exclusive scan over the P x ncols flags gives every
// partition's ranks at once (offset by the partition's first entry).
__device__ __forceinline__ int64_t part_of(const int2 *coords, int64_t P, int64_t j)
{
    int64_t lo = 0, hi = P;  // largest t in [0, P) with coords[t].y <= j
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (coords[mid].y <= j) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void k_ecr_mark(const int32_t *__restrict__ col, int64_t nnz, int64_t ncols,
                           const int2 *__restrict__ coords, int64_t P, int32_t *flags)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
         j += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = part_of(coords, P, j);
        flags[t * ncols + col[j]] = 1;
    }
}

__global__ void k_ecr_assign(const int32_t *__restrict__ col, int64_t nnz, int64_t ncols,
                             const int2 *__restrict__ coords, int64_t P,
                             const int32_t *__restrict__ rank, int32_t *__restrict__ ecr)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
         j += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = part_of(coords, P, j);
        ecr[j] = rank[t * ncols + col[j]] - rank[t * ncols];
    }
}

__global__ void k_ecr_win(int64_t total, int64_t ncols, const int32_t *__restrict__ flags,
                          const int32_t *__restrict__ rank, int32_t *__restrict__ win)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x)
        if (flags[i]) win[rank[i]] = (int32_t)(i % ncols);
}

__global__ void k_ecr_off(int64_t P, int64_t ncols, const int32_t *__restrict__ rank,
                          int64_t grand_total, int64_t *__restrict__ win_off)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < P) win_off[t] = rank[t * ncols];
    if (t == P) win_off[P] = grand_total;
}

extern "C" int vnec_hip_ecr_dense(const int32_t *col, int64_t nnz, int64_t ncols,
                                  const int32_t *coords, int64_t P, int32_t *flags, int32_t *rank,
                                  int32_t *ecr, int32_t *win, int64_t *win_off, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    int64_t total = P * ncols;
    HIPCHK(hipMemsetAsync(flags, 0, sizeof(int32_t) * (size_t)total, s));
    const int2 *c2 = (const int2 *)coords;
    unsigned g1 = (unsigned)((nnz + 255) / 256 < 65536 ? (nnz + 255) / 256 : 65536);
    if (g1 == 0) g1 = 1;
    hipLaunchKernelGGL(k_ecr_mark, dim3(g1), dim3(256), 0, s, col, nnz, ncols, c2, P, flags);
    HIPCHK(hipGetLastError());
    int64_t grand = 0;
    int rc = vnec_hip_scan_excl_i32(flags, rank, total, &grand, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ecr_assign, dim3(g1), dim3(256), 0, s, col, nnz, ncols, c2, P, rank, ecr);
    HIPCHK(hipGetLastError());
    unsigned g2 = (unsigned)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
    hipLaunchKernelGGL(k_ecr_win, dim3(g2), dim3(256), 0, s, total, ncols, flags, rank, win);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_ecr_off, dim3((unsigned)((P + 1 + 255) / 256)), dim3(256), 0, s, P, ncols,
                       rank, grand, win_off);
    return (int)hipGetLastError();
}

// --------------------------------------------- variant C: column-sorted tiles
// keys[j] = tile << cb | col[j] (cb bits hold any column), idx[j] =
// j - coords[tile].y (one workgroup per tile).
__global__ void k_csort_keys(const int2 *__restrict__ co, const int32_t *__restrict__ col, int cb,
                             uint64_t *__restrict__ keys, int32_t *__restrict__ idx)
{
    const int64_t t = blockIdx.x;
    const int y0 = co[t].y, y1 = co[t + 1].y;
    for (int j = y0 + (int)threadIdx.x; j < y1; j += blockDim.x) {
        keys[j] = ((uint64_t)t << cb) | (uint32_t)col[j];
        idx[j] = j - y0;
    }
}

// After the sort: position j holds tile t's nonzero at in-tile offset idx[j].
__global__ void k_csort_apply(const int2 *__restrict__ co, const uint64_t *__restrict__ keys,
                              const int32_t *__restrict__ idx, const double *__restrict__ val,
                              int64_t nnz, int cb, int pos_shift, const int32_t *__restrict__ cbase,
                              double *__restrict__ sval, int32_t *__restrict__ scol,
                              uint16_t *__restrict__ spos)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz;
         j += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[j];
        const int64_t t = (int64_t)(k >> cb);
        const int off = idx[j];
        sval[j] = val[co[t].y + off];
        const uint32_t c = (uint32_t)(k & (((uint64_t)1 << cb) - 1));
        if (pos_shift > 0) {
            const uint32_t cr = cbase ? c - (uint32_t)cbase[t] : c;
            scol[j] = (int32_t)(cr | ((uint32_t)off << pos_shift));
        } else {
            scol[j] = (int32_t)c;
            spos[j] = (uint16_t)off;
        }
    }
}

extern "C" int vnec_hip_csort(const int32_t *coords, int64_t ntiles, const int32_t *col, const double *val,
                              int64_t nnz, int64_t ncols, int pos_shift, const int32_t *cbase,
                              uint64_t *keys, int32_t *idx, double *sval, int32_t *scol, uint16_t *spos,
                              void *stream)
{
    if (nnz <= 0 || ntiles <= 0) return 0;
    int cb = 1, tb = 1;  // bits for a column, bits for a tile index
    while (cb < 31 && ((int64_t)1 << cb) < ncols) cb++;
    while (tb < 32 && ((int64_t)1 << tb) < ntiles) tb++;
    if (cb + tb > 64 || ntiles >= ((int64_t)1 << 31)) return (int)hipErrorInvalidValue;
    if (pos_shift < 0 || pos_shift > 31 || (pos_shift > 0 && !cbase && ncols > ((int64_t)1 << pos_shift)))
        return (int)hipErrorInvalidValue;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_csort_keys, dim3((unsigned)ntiles), dim3(256), 0, s, (const int2 *)coords, col, cb,
                       keys, idx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    int rc = vnec_hip_sort_pairs_u64(keys, idx, nnz, cb + tb, stream);
    if (rc) return rc;
    unsigned g = (unsigned)((nnz + 255) / 256 < 65536 ? (nnz + 255) / 256 : 65536);
    hipLaunchKernelGGL(k_csort_apply, dim3(g), dim3(256), 0, s, (const int2 *)coords, keys, idx, val, nnz, cb,
                       pos_shift, cbase, sval, scol, spos);
    return (int)hipGetLastError();
}

// One workgroup per tile: its smallest column, and the widest column span
// of any tile (atomicMax into *span).
__global__ __launch_bounds__(VNEC_BLOCK) void k_tile_colspan(const int2 *__restrict__ co,
                                                             const int32_t *__restrict__ col,
                                                             int32_t *__restrict__ cmin,
                                                             int32_t *__restrict__ span)
{
    __shared__ int s_mn[VNEC_NWAVES], s_mx[VNEC_NWAVES];
    const int64_t t = blockIdx.x;
    const int y0 = co[t].y, y1 = co[t + 1].y;
    int mn = INT_MAX, mx = -1;
    for (int j = y0 + (int)threadIdx.x; j < y1; j += VNEC_BLOCK) {
        const int c = col[j];
        mn = c < mn ? c : mn;
        mx = c > mx ? c : mx;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int a = __shfl_down(mn, o, VNEC_WAVE), b = __shfl_down(mx, o, VNEC_WAVE);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    const int w = threadIdx.x / VNEC_WAVE;
    if (lane_id() == 0) {
        s_mn[w] = mn;
        s_mx[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < VNEC_NWAVES; i++) {
            mn = s_mn[i] < mn ? s_mn[i] : mn;
            mx = s_mx[i] > mx ? s_mx[i] : mx;
        }
        const bool any = y1 > y0;
        cmin[t] = any ? mn : 0;
        if (any) atomicMax(span, mx - mn + 1);
    }
}

extern "C" int vnec_hip_tile_colspan(const int32_t *coords, int64_t ntiles, const int32_t *col, int32_t *cmin,
                                     int32_t *span_host, void *stream)
{
    if (ntiles <= 0 || ntiles >= ((int64_t)1 << 31) || !span_host) return (int)hipErrorInvalidValue;
    hipStream_t s = (hipStream_t)stream;
    int32_t *d = nullptr;
    HIPCHK(hipMalloc(&d, 4));
    int rc = (int)hipMemsetAsync(d, 0, 4, s);
    if (!rc) {
        hipLaunchKernelGGL(k_tile_colspan, dim3((unsigned)ntiles), dim3(VNEC_BLOCK), 0, s, (const int2 *)coords,
                           col, cmin, d);
        rc = (int)hipGetLastError();
    }
    if (!rc) rc = (int)hipMemcpyAsync(span_host, d, 4, hipMemcpyDeviceToHost, s);
    if (!rc) rc = (int)hipStreamSynchronize(s);
    hipError_t f = hipFree(d);
    return rc ? rc : (int)f;
}

// ------------------------------------------- right-hand-side layouts
// Column-major k x n (k vectors of n, one after another) <-> interleaved
// n x k (entry j of every vector together), k <= 16.  A workgroup moves 256
// entries of each vector through LDS, so both sides are read and written
// in contiguous runs.
#define XPOSE_N 256
__global__ __launch_bounds__(XPOSE_N) void k_to_interleaved(const double *__restrict__ src, int64_t n, int k,
                                                          double *__restrict__ dst)
{
    __shared__ double s[16][XPOSE_N + 1];
    const int64_t j0 = (int64_t)blockIdx.x * XPOSE_N;
    const int t = threadIdx.x;
    for (int c = 0; c < k; c++)
        if (j0 + t < n) s[c][t] = src[c * n + j0 + t];
    __syncthreads();
    const int64_t m = (n - j0 < XPOSE_N ? n - j0 : XPOSE_N) * k;
    for (int64_t e = t; e < m; e += XPOSE_N) dst[j0 * k + e] = s[e % k][e / k];
}

__global__ __launch_bounds__(XPOSE_N) void k_from_interleaved(const double *__restrict__ src, int64_t n, int k,
                                                            double *__restrict__ dst)
{
    __shared__ double s[16][XPOSE_N + 1];
    const int64_t j0 = (int64_t)blockIdx.x * XPOSE_N;
    const int t = threadIdx.x;
    const int64_t m = (n - j0 < XPOSE_N ? n - j0 : XPOSE_N) * k;
    for (int64_t e = t; e < m; e += XPOSE_N) s[e % k][e / k] = src[j0 * k + e];
    __syncthreads();
    for (int c = 0; c < k; c++)
        if (j0 + t < n) dst[c * n + j0 + t] = s[c][t];
}

extern "C" int vnec_hip_rhs_layout(const double *src, int64_t n, int k, int to_interleaved, double *dst,
                                   void *stream)
{
    if (n <= 0) return 0;
    if (k < 1 || k > 16) return (int)hipErrorInvalidValue;
    const int64_t g = (n + XPOSE_N - 1) / XPOSE_N;
    if (g >= ((int64_t)1 << 31)) return (int)hipErrorInvalidValue;
    if (to_interleaved)
        hipLaunchKernelGGL(k_to_interleaved, dim3((unsigned)g), dim3(XPOSE_N), 0, (hipStream_t)stream, src, n, k,
                           dst);
    else
        hipLaunchKernelGGL(k_from_interleaved, dim3((unsigned)g), dim3(XPOSE_N), 0, (hipStream_t)stream, src, n, k,
                           dst);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------- value split
__global__ void k_split_hilo(const double *__restrict__ v, int64_t n, float *__restrict__ hi,
                             float *__restrict__ lo)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double a = v[i];
        float h = (float)a;
        hi[i] = h;
        if (lo) lo[i] = (float)(a - (double)h);
    }
}

extern "C" int vnec_hip_split_hilo(const double *val, int64_t n, float *hi, float *lo,
                                   void *stream)
{
    if (n <= 0) return 0;
    unsigned g = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    hipLaunchKernelGGL(k_split_hilo, dim3(g), dim3(256), 0, (hipStream_t)stream, val, n, hi, lo);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------- scribble
__global__ void k_scribble(uint4 *p, int64_t n16, uint32_t salt)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4((uint32_t)i ^ salt, salt, (uint32_t)(i >> 32), ~salt);
}

// y[0 .. n) = 0 with a kernel (also inside captured graphs, where memset
// nodes are avoided).
__global__ void k_zero_f64(double *__restrict__ y, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = 0.0;
}

extern "C" int vnec_hip_zero_f64(double *y, int64_t n, void *stream)
{
    if (n <= 0) return 0;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, y, n);
    return (int)hipGetLastError();
}

__global__ void k_zero_f32(float *__restrict__ y, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = 0.0f;
}

extern "C" int vnec_hip_zero_f32(float *y, int64_t n, void *stream)
{
    if (n <= 0) return 0;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, y, n);
    return (int)hipGetLastError();
}

// Device mirror of an FP32 matrix: its values widened (exactly) to FP64
// for the layout builders, which are shared with FP64 matrices.
__global__ void k_widen_f32(const float *__restrict__ in, int64_t n, double *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

extern "C" int vnec_hip_widen_f32(const float *in, int64_t n, double *out, void *stream)
{
    if (n <= 0) return 0;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_widen_f32, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, in, n, out);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_scribble(void *p, size_t n, void *stream)
{
    static uint32_t salt = 0x9E3779B9u;
    salt = salt * 1664525u + 1013904223u;
    int64_t n16 = (int64_t)(n / 16);
    if (n16 <= 0) return 0;
    hipLaunchKernelGGL(k_scribble, dim3(4096), dim3(256), 0, (hipStream_t)stream, (uint4 *)p, n16,
                       salt);
    return (int)hipGetLastError();
}

// ------------------------------------------------------------ copy probe
// Streaming copy of n16 16-byte words: the measured HBM ceiling that the
// SpMV kernels are compared with, and the calibration kernel for the
// hardware byte counters (read n bytes, write n bytes per launch).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Each thread keeps STREAM_U 16-byte loads in flight before it stores them.
#define STREAM_U 4

__global__ __launch_bounds__(256) void k_stream_copy(const u32x4 *__restrict__ src,
                                                     u32x4 *__restrict__ dst, int64_t n16)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (STREAM_U - 1) * stride < n16; i += STREAM_U * stride) {
        u32x4 v[STREAM_U];
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) __builtin_nontemporal_store(v[u], dst + i + u * stride);
    }
    for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

extern "C" int vnec_hip_stream_copy(const void *src, void *dst, size_t bytes, void *stream)
{
    int64_t n16 = (int64_t)(bytes / 16);
    if (n16 <= 0) return 0;
    hipLaunchKernelGGL(k_stream_copy, dim3(4096), dim3(256), 0, (hipStream_t)stream,
                       (const u32x4 *)src, (u32x4 *)dst, n16);
    return (int)hipGetLastError();
}

// Streaming read: each workgroup reads one contiguous 64 KiB chunk with
// 8-byte nontemporal loads, READ_U in flight per thread, folded into one
// word per thread that is stored only if it equals a value the probe's
// input never produces.  Contiguous chunks per workgroup read faster than a
// grid-strided sweep (7.0-7.4 against 6.5 TB/s measured on MI355X): it is
// the read-only HBM ceiling the SpMV kernels are compared with (an SpMV
// reads far more than it writes).
#define READ_U 8
#define READ_CHUNK_WORDS 8192 /* 64 KiB of 8-byte words per workgroup */

template <bool NT>
__global__ __launch_bounds__(256) void k_stream_read(const uint64_t *__restrict__ src, int64_t n8, uint32_t key,
                                                     uint32_t *__restrict__ out)
{
    const int64_t beg = (int64_t)blockIdx.x * READ_CHUNK_WORDS;
    const int64_t end = beg + READ_CHUNK_WORDS < n8 ? beg + READ_CHUNK_WORDS : n8;
    uint64_t acc = 0;
    int64_t i = beg + threadIdx.x;
    for (; i + (READ_U - 1) * 256 < end; i += READ_U * 256) {
        uint64_t v[READ_U];
#pragma unroll
        for (int u = 0; u < READ_U; u++) v[u] = NT ? __builtin_nontemporal_load(src + i + u * 256) : src[i + u * 256];
#pragma unroll
        for (int u = 0; u < READ_U; u++) acc ^= v[u];
    }
    for (; i < end; i += 256) acc ^= NT ? __builtin_nontemporal_load(src + i) : src[i];
    const uint32_t r = (uint32_t)acc ^ (uint32_t)(acc >> 32);
    if (r == key) out[blockIdx.x & 4095] = r;
}

static int stream_read(const void *src, size_t bytes, void *out, bool nt, void *stream)
{
    int64_t n8 = (int64_t)(bytes / 8);
    if (n8 <= 0) return 0;
    const int64_t blocks = (n8 + READ_CHUNK_WORDS - 1) / READ_CHUNK_WORDS;
    if (blocks > INT32_MAX) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(nt ? k_stream_read<true> : k_stream_read<false>, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, (const uint64_t *)src, n8, 0x9E3779B9u, (uint32_t *)out);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_stream_read(const void *src, size_t bytes, void *out, void *stream)
{
    return stream_read(src, bytes, out, true, stream);
}

// The same read with plain loads, which allocate in L2 and the Infinity
// Cache: reading a buffer larger than both evicts what an SpMV left there,
// without leaving dirty lines behind (a cache flush between timed SpMVs).
extern "C" int vnec_hip_flush_read(const void *src, size_t bytes, void *out, void *stream)
{
    return stream_read(src, bytes, out, false, stream);
}

// ----------------------------------------------------- gather ceiling
// What one SpMV's memory traffic costs without any row work: stream a
// matrix's values and columns (and, with GATHER, gather x[col] for each
// nonzero) over a persistent grid, 8 nonzeros per thread per chunk, one
// partial sum per thread written.  Scattered patterns such as R-MAT are
// bound by this gather rate rather than by HBM.
template <bool GATHER>
__global__ __launch_bounds__(256) void k_gather_probe(const double *__restrict__ val,
                                                      const int32_t *__restrict__ col, int64_t nnz,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ out)
{
    constexpr int U = 8;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t nchunks = (nnz + 256 * U - 1) / (256 * U);
    double s = 0.0;
    for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const int64_t base = (ch * 4 + w) * 64 * U + lane;
        double v[U];
        int c[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t j = base + 64 * u;
            v[u] = j < nnz ? __builtin_nontemporal_load(val + j) : 0.0;
            c[u] = j < nnz ? __builtin_nontemporal_load(col + j) : 0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) s += v[u] * (GATHER ? x[c[u]] : (double)c[u]);
    }
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

extern "C" int vnec_hip_gather_probe(const double *val, const int32_t *col, int64_t nnz, const double *x,
                                     int gather, double *out, int grid, void *stream)
{
    if (grid < 1) return (int)hipErrorInvalidValue;
    if (gather)
        hipLaunchKernelGGL(k_gather_probe<true>, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream,
                           val, col, nnz, x, out);
    else
        hipLaunchKernelGGL(k_gather_probe<false>, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream,
                           val, col, nnz, x, out);
    return (int)hipGetLastError();
}
