// vnec_sorder.hip — GPU builder of the VNEC-S task order and reordered
// matrix (spmv_vnec_S_for_mxv.c:442-585) at any (P, L, cross-row cut).
//
// Per snapped partition the reference buckets rows by length (ascending,
// ascending row id within a length), sends the first cnt - cnt % L rows of
// each length <= 32 to the cross-row part and the rest, with every longer
// row, to the in-row part, concatenates the two, copies the rows in that
// order and rewrites each cross-row group of L equal-length rows from
// [row][l] to [l][row].  Here that is two stable sorts:
//   1. (partition, length) over the rows in row order -> rank k of a row
//      among the rows of its length in its partition, and their count c;
//      class 0 (cross-row) when length <= cut and k < c - c % L, else 1;
//   2. (partition, class) over the result (stable, so length and row
//      order survive) -> the task list;
// followed by one wave per task row that copies its values and window
// ranks, transposing cross-row groups on the fly.  The sorts are a
// hand-written LSD radix sort with 8-bit digits whose scatter ranks equal
// digits inside a wavefront with ballots (stable by construction).
#include <hip/hip_runtime.h>

#include "vnec_common.h"
#include "vnec_hip.h"

#define RS_ROUNDS 16
#define RS_ITEMS (VNEC_BLOCK * RS_ROUNDS)

__global__ __launch_bounds__(VNEC_BLOCK) void k_radix_hist(const uint64_t *__restrict__ keys,
                                                           int64_t n, int shift, int64_t nblocks,
                                                           int32_t *__restrict__ counts)
{
    __shared__ int s_hist[256];
    s_hist[threadIdx.x] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * RS_ITEMS;
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int64_t i = b0 + r * VNEC_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&s_hist[(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    counts[(int64_t)threadIdx.x * nblocks + blockIdx.x] = s_hist[threadIdx.x];
}

__global__ __launch_bounds__(VNEC_BLOCK) void k_radix_scatter(
    const uint64_t *__restrict__ kin, const int32_t *__restrict__ vin, int64_t n, int shift,
    int64_t nblocks, const int32_t *__restrict__ offsets, uint64_t *__restrict__ kout,
    int32_t *__restrict__ vout)
{
    __shared__ int s_off[256];
    __shared__ int s_wcnt[VNEC_NWAVES][256];
    const int tid = threadIdx.x, w = wave_id(), lane = lane_id();
    s_off[tid] = offsets[(int64_t)tid * nblocks + blockIdx.x];
    for (int k = 0; k < VNEC_NWAVES; k++) s_wcnt[k][tid] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * RS_ITEMS;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int64_t i = b0 + r * VNEC_BLOCK + tid;
        const bool valid = i < n;
        uint64_t key = 0;
        int32_t v = 0;
        int d = 0;
        if (valid) {
            key = kin[i];
            v = vin[i];
            d = (int)((key >> shift) & 255u);
        }
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const bool set = (d >> bit) & 1;
            const uint64_t bb = __ballot(valid && set);
            peers &= set ? bb : ~bb;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) s_wcnt[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = s_off[d] + rank;
            for (int k = 0; k < w; k++) pos += s_wcnt[k][d];
            kout[pos] = key;
            vout[pos] = v;
        }
        __syncthreads();
        int add = 0;
        for (int k = 0; k < VNEC_NWAVES; k++) {
            add += s_wcnt[k][tid];
            s_wcnt[k][tid] = 0;
        }
        s_off[tid] += add;
        __syncthreads();
    }
}

// Stable sort of (key, value) pairs by the low key_bits bits of the keys.
// Allocates its own scratch (builder use only).  The sorted pairs end in
// keys / vals.
extern "C" int vnec_hip_sort_pairs_u64(uint64_t *keys, int32_t *vals, int64_t n, int key_bits,
                                       void *stream)
{
    if (n <= 1) return 0;
    if (n >= INT32_MAX) return (int)hipErrorInvalidValue;
    hipStream_t s = (hipStream_t)stream;
    const int64_t nblocks = (n + RS_ITEMS - 1) / RS_ITEMS;
    uint64_t *k2 = nullptr;
    int32_t *v2 = nullptr, *counts = nullptr;
    int rc = (int)hipMalloc((void **)&k2, sizeof(uint64_t) * (size_t)n);
    if (!rc) rc = (int)hipMalloc((void **)&v2, sizeof(int32_t) * (size_t)n);
    if (!rc) rc = (int)hipMalloc((void **)&counts, sizeof(int32_t) * (size_t)(256 * nblocks));
    uint64_t *ka = keys, *kb = k2;
    int32_t *va = vals, *vb = v2;
    for (int shift = 0; !rc && shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nblocks), dim3(VNEC_BLOCK), 0, s, ka, n, shift,
                           nblocks, counts);
        rc = (int)hipGetLastError();
        if (!rc) rc = vnec_hip_scan_excl_i32(counts, counts, 256 * nblocks, nullptr, stream);
        if (!rc) {
            hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nblocks), dim3(VNEC_BLOCK), 0, s, ka, va,
                               n, shift, nblocks, counts, kb, vb);
            rc = (int)hipGetLastError();
        }
        uint64_t *tk = ka;
        ka = kb;
        kb = tk;
        int32_t *tv = va;
        va = vb;
        vb = tv;
    }
    if (!rc && ka != keys) rc = (int)hipMemcpyAsync(keys, ka, sizeof(uint64_t) * (size_t)n, hipMemcpyDeviceToDevice, s);
    if (!rc && va != vals) rc = (int)hipMemcpyAsync(vals, va, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice, s);
    hipError_t e = hipStreamSynchronize(s);
    if (!rc) rc = (int)e;
    if (k2) (void)hipFree(k2);
    if (v2) (void)hipFree(v2);
    if (counts) (void)hipFree(counts);
    return rc;
}

// Partition of row r: largest t < P with start_x[t] <= r (start_x = the
// snapped start rows, coords[2t]).
__device__ __forceinline__ int64_t part_of_row(const int2 *co, int64_t P, int64_t r)
{
    int64_t lo = 0, hi = P;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (co[mid].x <= r) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void k_s_keys(const int32_t *__restrict__ ptr, int64_t nrows, const int2 *__restrict__ co,
                         int64_t P, uint64_t *__restrict__ keys, int32_t *__restrict__ vals)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
         r += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t t = (uint64_t)part_of_row(co, P, r);
        keys[r] = (t << 32) | (uint32_t)(ptr[r + 1] - ptr[r]);
        vals[r] = (int32_t)r;
    }
}

__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t *a, int64_t n, uint64_t v)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// keys sorted by (partition, length): class of each row, new key
// (partition << 1 | class) written to keys2.
__global__ void k_s_class(const uint64_t *__restrict__ keys, int64_t n, int L, int cut,
                          uint64_t *__restrict__ keys2)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = keys[i];
        const int64_t lb = lower_bound_u64(keys, n, key);
        const int64_t ub = lower_bound_u64(keys, n, key + 1);
        const int64_t k = i - lb, c = ub - lb;
        const int64_t len = (int64_t)(key & 0xFFFFFFFFull);
        const uint64_t cls = (len <= cut && k < c - c % L) ? 0u : 1u;
        keys2[i] = ((key >> 32) << 1) | cls;
    }
}

__global__ void k_s_spvv(const uint64_t *__restrict__ keys2, int64_t n, int64_t P,
                         int32_t *__restrict__ spvv)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P) return;
    const uint64_t k0 = (uint64_t)t << 1;
    spvv[t] = (int32_t)(lower_bound_u64(keys2, n, k0 + 1) - lower_bound_u64(keys2, n, k0));
}

extern "C" int vnec_hip_s_order(const int32_t *ptr, int64_t nrows, const int32_t *coords, int64_t P,
                                int L, int cut, int32_t *tasks, int32_t *spvv_len, void *stream)
{
    if (nrows <= 0 || P <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    uint64_t *keys = nullptr;
    int rc = (int)hipMalloc((void **)&keys, sizeof(uint64_t) * (size_t)nrows);
    int pbits = 1;
    while (pbits < 62 && (1ll << pbits) < P) pbits++;
    const int64_t blocks = (nrows + 255) / 256 < 65536 ? (nrows + 255) / 256 : 65536;
    if (!rc) {
        hipLaunchKernelGGL(k_s_keys, dim3((unsigned)blocks), dim3(256), 0, s, ptr, nrows,
                           (const int2 *)coords, P, keys, tasks);
        rc = (int)hipGetLastError();
    }
    if (!rc) rc = vnec_hip_sort_pairs_u64(keys, tasks, nrows, 32 + pbits, stream);
    uint64_t *keys2 = nullptr;
    if (!rc) rc = (int)hipMalloc((void **)&keys2, sizeof(uint64_t) * (size_t)nrows);
    if (!rc) {
        hipLaunchKernelGGL(k_s_class, dim3((unsigned)blocks), dim3(256), 0, s, keys, nrows, L, cut, keys2);
        rc = (int)hipGetLastError();
    }
    if (!rc) rc = vnec_hip_sort_pairs_u64(keys2, tasks, nrows, pbits + 1, stream);
    if (!rc) {
        hipLaunchKernelGGL(k_s_spvv, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, keys2, nrows,
                           P, spvv_len);
        rc = (int)hipGetLastError();
    }
    hipError_t e = hipStreamSynchronize(s);
    if (!rc) rc = (int)e;
    if (keys) (void)hipFree(keys);
    if (keys2) (void)hipFree(keys2);
    return rc;
}

__global__ void k_s_lens(const int32_t *__restrict__ ptr, const int32_t *__restrict__ tasks,
                         int64_t n, int32_t *__restrict__ lens)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n;
         i += (int64_t)gridDim.x * blockDim.x)
        lens[i] = i < n ? ptr[tasks[i] + 1] - ptr[tasks[i]] : 0;
}

__global__ void k_s_add(int32_t *__restrict__ a, const int32_t *__restrict__ b, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] += b[i];
}

// One wave per task position: copy the row's values and ECR ranks to its
// place, transposed to [l][row] inside cross-row groups.
__global__ __launch_bounds__(VNEC_BLOCK) void k_s_reorder(
    const int32_t *__restrict__ ptr, const double *__restrict__ val, const int32_t *__restrict__ ecr,
    const int32_t *__restrict__ tasks, int64_t n, const int32_t *__restrict__ row_begin,
    const int2 *__restrict__ co, const int32_t *__restrict__ spvv, int64_t P, int L, int64_t nnz,
    double *__restrict__ s_val, int32_t *__restrict__ s_col)
{
    const int64_t nw = (int64_t)gridDim.x * VNEC_NWAVES;
    for (int64_t pos = (int64_t)blockIdx.x * VNEC_NWAVES + wave_id(); pos < n; pos += nw) {
        const int64_t t = part_of_row(co, P, pos);  // task positions follow the snapped rows
        const int64_t ts = co[t].x;
        const int64_t k = pos - ts;
        const int32_t row = tasks[pos];
        const int32_t b = ptr[row], len = ptr[row + 1] - b;
        int64_t base, stride;
        if (k < spvv[t]) {  // cross-row: element l of lane r at base + l * L + r
            const int64_t c0 = ts + (k / L) * L;
            base = (int64_t)row_begin[c0] + (k % L);
            stride = L;
        } else {
            base = row_begin[pos];
            stride = 1;
        }
        for (int l = lane_id(); l < len; l += VNEC_WAVE) {
            const int64_t o = base + (int64_t)l * stride;
            if (o >= nnz) break;  // only reachable if a cross-row group mixed lengths
            if (s_val) s_val[o] = val[b + l];
            s_col[o] = ecr[b + l];
        }
    }
}

extern "C" int vnec_hip_s_reorder(const int32_t *ptr, const double *val, const int32_t *ecr,
                                  const int32_t *tasks, int64_t nrows, int64_t nnz, const int32_t *coords,
                                  const int32_t *spvv_len, int64_t P, int L, int32_t *row_begin,
                                  int32_t *row_end, double *s_val, int32_t *s_col, void *stream)
{
    if (nrows <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    int32_t *lens = nullptr;
    int rc = (int)hipMalloc((void **)&lens, sizeof(int32_t) * (size_t)(nrows + 1));
    const int64_t blocks = (nrows + 256) / 256 < 65536 ? (nrows + 256) / 256 : 65536;
    if (!rc) {
        hipLaunchKernelGGL(k_s_lens, dim3((unsigned)blocks), dim3(256), 0, s, ptr, tasks, nrows, lens);
        rc = (int)hipGetLastError();
    }
    if (!rc) rc = (int)hipMemcpyAsync(row_end, lens, sizeof(int32_t) * (size_t)nrows, hipMemcpyDeviceToDevice, s);
    if (!rc) rc = vnec_hip_scan_excl_i32(lens, lens, nrows + 1, nullptr, stream);
    if (!rc) rc = (int)hipMemcpyAsync(row_begin, lens, sizeof(int32_t) * (size_t)nrows, hipMemcpyDeviceToDevice, s);
    if (!rc) {
        // row_end = row_begin + length
        hipLaunchKernelGGL(k_s_add, dim3((unsigned)blocks), dim3(256), 0, s, row_end, row_begin, nrows);
        rc = (int)hipGetLastError();
    }
    if (!rc) {
        const int64_t wblocks = (nrows + VNEC_NWAVES - 1) / VNEC_NWAVES < 16384 ? (nrows + VNEC_NWAVES - 1) / VNEC_NWAVES : 16384;
        hipLaunchKernelGGL(k_s_reorder, dim3((unsigned)wblocks), dim3(VNEC_BLOCK), 0, s, ptr, val, ecr, tasks,
                           nrows, row_begin, (const int2 *)coords, spvv_len, P, L, nnz, s_val, s_col);
        rc = (int)hipGetLastError();
    }
    hipError_t e = hipStreamSynchronize(s);
    if (!rc) rc = (int)e;
    if (lens) (void)hipFree(lens);
    return rc;
}
