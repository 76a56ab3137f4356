// vnec_layout.hip — GPU builder for the VNEC-D block layout
// (spmv_vnec_D_for_mxv.c:356-392) at any (P, L).
//
// Every row belongs to the partition that completes it (coords[t].x <= i
// < coords[t + 1].x); its segment runs from max(ptr[i], coords[t].y) to
// ptr[i + 1].  The greedy grouping is row-local, so rows are processed in
// parallel: pass 1 counts each row's groups, a scan gives v_row_ptr, and
// pass 2 writes col_start and the zero-padded val_align.
#include <hip/hip_runtime.h>

#include "vnec_common.h"
#include "vnec_hip.h"

__device__ __forceinline__ int64_t owner_of_row(const int2 *co, int64_t P, int64_t i)
{
    int64_t lo = 0, hi = P;  // largest t in [0, P) with co[t].x <= i
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (co[mid].x <= i) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void k_d_count(const int32_t *__restrict__ ptr, const int32_t *__restrict__ ecr,
                          int64_t nrows, const int2 *__restrict__ co, int64_t P, int L,
                          int32_t *__restrict__ cnt, int32_t *__restrict__ unsorted)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= nrows;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i == nrows) {
            cnt[i] = 0;
            continue;
        }
        int64_t t = owner_of_row(co, P, i);
        int64_t b = ptr[i] > co[t].y ? ptr[i] : co[t].y;
        int64_t e = ptr[i + 1];
        int g = 0, desc = 0;
        if (e > b) {
            int c0 = ecr[b], prev = c0;
            g = 1;
            for (int64_t j = b + 1; j < e; j++) {
                int c = ecr[j];
                desc |= c < prev;
                prev = c;
                if (c - c0 >= L) {
                    g++;
                    c0 = c;
                }
            }
        }
        cnt[i] = g;
        if (desc && unsorted) atomicOr(unsorted, 1);
    }
}

__global__ void k_d_fill(const int32_t *__restrict__ ptr, const int32_t *__restrict__ ecr,
                         const double *__restrict__ val, int64_t nrows,
                         const int2 *__restrict__ co, int64_t P, int L,
                         const int32_t *__restrict__ vrp, int32_t *__restrict__ col_start,
                         double *__restrict__ val_align, int accumulate)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = owner_of_row(co, P, i);
        int64_t b = ptr[i] > co[t].y ? ptr[i] : co[t].y;
        int64_t e = ptr[i + 1];
        if (e <= b) continue;
        int64_t g = vrp[i];
        int c0 = ecr[b];
        col_start[g] = c0;
        if (val_align) val_align[g * L] = val[b];
        for (int64_t j = b + 1; j < e; j++) {
            int c = ecr[j];
            int dist = c - c0;
            if (dist < 0) continue;  // unsorted row: flagged by k_d_count, never written
            if (dist < L) {
                // a repeated column overwrites (the reference's layout) or adds
                if (val_align) val_align[g * L + dist] = accumulate ? val_align[g * L + dist] + val[j] : val[j];
            } else {
                g++;
                c0 = c;
                col_start[g] = c0;
                if (val_align) val_align[g * L] = val[j];
            }
        }
    }
}

extern "C" int vnec_hip_d_count(const int32_t *ptr, const int32_t *ecr, int64_t nrows,
                                const int32_t *coords, int64_t P, int L, int32_t *cnt,
                                int32_t *unsorted, void *stream)
{
    int64_t n = nrows + 1;
    unsigned g = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    hipLaunchKernelGGL(k_d_count, dim3(g), dim3(256), 0, (hipStream_t)stream, ptr, ecr, nrows,
                       (const int2 *)coords, P, L, cnt, unsorted);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_d_fill(const int32_t *ptr, const int32_t *ecr, const double *val,
                               int64_t nrows, const int32_t *coords, int64_t P, int L,
                               const int32_t *v_row_ptr, int32_t *col_start, double *val_align,
                               int accumulate, void *stream)
{
    if (nrows <= 0) return 0;
    unsigned g = (unsigned)((nrows + 255) / 256 < 65536 ? (nrows + 255) / 256 : 65536);
    hipLaunchKernelGGL(k_d_fill, dim3(g), dim3(256), 0, (hipStream_t)stream, ptr, ecr, val, nrows,
                       (const int2 *)coords, P, L, v_row_ptr, col_start, val_align, accumulate);
    return (int)hipGetLastError();
}
