// vnec_pattern.hip — variant G: pattern slices for gfx950.
//
// Many structured matrices (stencils, banded systems) have one dominant
// row pattern: a list of column offsets P[0 .. L) such that row i's entries
// sit at columns i + P[l] for a subset of l — all of them inside the
// domain, the ones that would fall outside dropped at its boundary.  For
// such a row the column indices carry nothing beyond the L-bit mask of
// which pattern positions it holds, so the plan stores the mask and the
// values only: 64 consecutive rows form a slice whose values are laid out
// [l][lane], zero where the mask bit is clear (one coalesced 512-byte load
// per wavefront per l).  A lane owns one row and reads x[row + P[l]] for
// the positions it holds: consecutive rows read consecutive x entries, so
// these gathers coalesce too.  Every row is summed over l, i.e. in its CSR
// column order.
//
// Rows that are not such a subset (a column outside the pattern, or one
// repeated) get mask 0 here — the kernel writes y = 0 for them — and form
// a CSR sub-matrix that an ordinary plan runs into a scratch vector,
// scattered back over those rows by k_scatter_rows afterwards.
//
// Per stored slot the kernel moves 8 bytes (the value; 4 for the FP32
// hi-only encoding, 4 + 4 for hi + lo) plus 4 bytes of mask per row,
// instead of 10-12 bytes per nonzero for a column-indexed format.
#include <hip/hip_runtime.h>

#include "vnec_common.h"
#include "vnec_hip.h"

// masks[i] = bits of the pattern positions row i holds, flags[i] = 1 when
// row i is not a subset of the pattern (its entries must then go through
// the sub-plan).  Columns ascend within a row, as do the offsets.
__global__ void k_pattern_match(const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                                int64_t nrows, const int32_t *__restrict__ pat, int L,
                                uint32_t *__restrict__ masks, int32_t *__restrict__ flags)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int b = ptr[i], e = ptr[i + 1];
        uint32_t m = 0;
        int ok = e - b <= L, l = 0;
        for (int j = b; ok && j < e; j++) {
            const int64_t off = (int64_t)col[j] - i;
            while (l < L && (int64_t)pat[l] < off) l++;
            if (l < L && (int64_t)pat[l] == off) {
                m |= 1u << l;
                l++;
            } else {
                ok = 0;
            }
        }
        masks[i] = ok ? m : 0u;
        flags[i] = !ok;
    }
}

// pos = exclusive scan of flags: irregular row i goes to irr[pos[i]].
__global__ void k_pattern_split(const int32_t *__restrict__ flags, const int32_t *__restrict__ pos,
                                int64_t nrows, int32_t *__restrict__ irr)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows;
         i += (int64_t)gridDim.x * blockDim.x)
        if (flags[i]) irr[pos[i]] = (int32_t)i;
}

// gv[(s L + l) 64 + lane] = value of row 64 s + lane at pattern position l
// (its k-th stored entry, k = set bits of the mask below l), else 0.
__global__ void k_pattern_fill(const int32_t *__restrict__ ptr, const double *__restrict__ val,
                               const uint32_t *__restrict__ masks, int64_t nrows, int64_t nslices,
                               int L, double *__restrict__ gv)
{
    const int64_t n = nslices * L * 64;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = t / (64 * (int64_t)L), r = t - s * 64 * L;
        const int l = (int)(r / 64), lane = (int)(r % 64);
        const int64_t row = s * 64 + lane;
        double v = 0.0;
        if (row < nrows) {
            const uint32_t m = masks[row];
            if ((m >> l) & 1u) v = val[ptr[row] + __popc(m & ((1u << l) - 1u))];
        }
        gv[t] = v;
    }
}

// Slot value k of the slices in the kernel's arithmetic: FP64 from stored
// FP64, FP32 hi + lo or hi; FP32 for the FP32 API (encoding 3).
template <int ENC>
__device__ __forceinline__ typename enc_real<ENC>::type slot_value(const vnec_pattern_args &a, int64_t k)
{
    if (ENC == 0) return __builtin_nontemporal_load(a.gv + k);
    if (ENC == 1)
        return (double)__builtin_nontemporal_load(a.gh + k) + (double)__builtin_nontemporal_load(a.gl + k);
    return __builtin_nontemporal_load(a.gh + k);  // hi (widened for ENC 2; FP32 for ENC 3)
}
__device__ __forceinline__ const double *pattern_x(const vnec_pattern_args &a, double) { return a.x; }
__device__ __forceinline__ const float *pattern_x(const vnec_pattern_args &a, float) { return a.x32; }
__device__ __forceinline__ double *pattern_y(const vnec_pattern_args &a, double) { return a.y; }
__device__ __forceinline__ float *pattern_y(const vnec_pattern_args &a, float) { return a.y32; }

// Rows (and threads) per workgroup of the pattern kernel: eight slices.
// A wider workgroup shares its x window's halo among more rows (27-point
// stencil: 4.5 staged entries per row at 512 rows, 6 at 256).
#define G_ROWS VNEC_PATTERN_ROWS
#define G_WAVES (G_ROWS / VNEC_WAVE)

// Stages the window's runs into LDS with 16-byte loads: run k holds
// columns r0 + wst[k] + [0, wlen[k]); each lane loads the 16-byte-aligned
// group of x entries that covers its share (groups are aligned to x's own
// address, whatever its alignment) and stores the entries that fall in the
// run; columns outside [0, ncols) stage as 0.
template <typename F>
__device__ __forceinline__ void stage_window(const vnec_pattern_args &a, const F *__restrict__ x, int64_t r0, F *s_xw)
{
    constexpr int E = 16 / sizeof(F);  // entries per 16-byte group
    typedef F vec __attribute__((ext_vector_type(E)));
    const int mis = (int)(((uintptr_t)x / sizeof(F)) % E);  // x + c is 16-byte aligned iff (c + mis) % E == 0
    for (int k = 0; k < a.nwin; k++) {
        const int64_t c0 = r0 + a.wst[k];
        const int64_t g0 = c0 - ((c0 + mis) % E + E) % E;  // aligned group start <= c0
        const int sh = (int)(c0 - g0);
        const int ng = (a.wlen[k] + sh + E - 1) / E;
        for (int i = threadIdx.x; i < ng; i += G_ROWS) {
            const int64_t c = g0 + (int64_t)i * E;
            F v[E];
            if (c >= 0 && c + E <= a.ncols) {
                const vec w = *(const vec *)(x + c);
#pragma unroll
                for (int e = 0; e < E; e++) v[e] = w[e];
            } else {
#pragma unroll
                for (int e = 0; e < E; e++) v[e] = c + e >= 0 && c + e < a.ncols ? x[c + e] : (F)0;
            }
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int p = i * E + e - sh;  // window position
                if (p >= 0 && p < a.wlen[k]) s_xw[a.wbase[k] + p] = v[e];
            }
        }
    }
}

// One wavefront per slice, G_WAVES slices per workgroup, XCD-aware order.
// WIN: the workgroup first stages its x window in LDS (a.wsize entries,
// 16-byte loads), so the rows' x reads, each entry shared by up to L rows
// of neighbouring slices, are LDS reads instead of L2 requests.
template <int U, int ENC, bool WIN>
__global__ __launch_bounds__(G_ROWS) void k_spmv_pattern(vnec_pattern_args a, int64_t grid)
{
    typedef typename enc_real<ENC>::type F;
    extern __shared__ double s_dyn[];
    F *s_xw = (F *)s_dyn;
    const F *__restrict__ x = pattern_x(a, F());
    const int64_t b = xcd_tile(blockIdx.x, grid);
    const int64_t r0 = b * G_ROWS;  // first row of the workgroup
    if (r0 >= a.nrows) return;       // whole workgroup
    if (WIN) {
        stage_window(a, x, r0, s_xw);
        __syncthreads();
    }
    const int64_t s = b * G_WAVES + wave_id();
    if (s >= a.nslices) return;
    const int t = (int)threadIdx.x;  // local row
    const int64_t row = r0 + t;
    const bool live = row < a.nrows;
    const uint32_t m = live ? __builtin_nontemporal_load(a.masks + row) : 0u;
    const int64_t v = s * (int64_t)a.L * 64 + lane_id();  // slot (slice s, l = 0, this lane)
    F acc = 0;
    int l = 0;
    for (; l + U <= a.L; l += U) {
        F vv[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) vv[u] = slot_value<ENC>(a, v + (int64_t)(l + u) * 64);
#pragma unroll
        for (int u = 0; u < U; u++)
            xv[u] = ((m >> (l + u)) & 1u) ? (WIN ? s_xw[a.wpos[l + u] + t] : x[row + a.pat[l + u]]) : (F)0;
#pragma unroll
        for (int u = 0; u < U; u++) acc += vv[u] * xv[u];
    }
    for (; l < a.L; l++) {
        const F vv = slot_value<ENC>(a, v + (int64_t)l * 64);
        if ((m >> l) & 1u) acc += vv * (WIN ? s_xw[a.wpos[l] + t] : x[row + a.pat[l]]);
    }
    if (live) __builtin_nontemporal_store(acc, pattern_y(a, F()) + row);  // y is not read back here
}

template <typename F>
__global__ void k_scatter_rows(const F *__restrict__ ysub, const int32_t *__restrict__ rows, int64_t n,
                               F *__restrict__ y)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[rows[i]] = ysub[i];
}

static unsigned grid_for(int64_t n)
{
    int64_t g = (n + 255) / 256;
    return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

extern "C" int vnec_hip_pattern_match(const int32_t *ptr, const int32_t *col, int64_t nrows,
                                      const int32_t *pat, int L, uint32_t *masks, int32_t *flags,
                                      void *stream)
{
    if (nrows <= 0) return 0;
    if (L < 1 || L > VNEC_PATTERN_MAX) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pattern_match, dim3(grid_for(nrows)), dim3(256), 0, (hipStream_t)stream,
                       ptr, col, nrows, pat, L, masks, flags);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_pattern_split(const int32_t *flags, const int32_t *pos, int64_t nrows,
                                      int32_t *irr, void *stream)
{
    if (nrows <= 0) return 0;
    hipLaunchKernelGGL(k_pattern_split, dim3(grid_for(nrows)), dim3(256), 0, (hipStream_t)stream,
                       flags, pos, nrows, irr);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_pattern_fill(const int32_t *ptr, const double *val, const uint32_t *masks,
                                     int64_t nrows, int L, double *gv, void *stream)
{
    const int64_t ns = (nrows + 63) / 64;
    if (ns <= 0) return 0;
    if (L < 1 || L > VNEC_PATTERN_MAX) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pattern_fill, dim3(grid_for(ns * L * 64)), dim3(256), 0, (hipStream_t)stream,
                       ptr, val, masks, nrows, ns, L, gv);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_spmv_pattern(const vnec_pattern_args *a, void *stream)
{
    if (a->nslices <= 0) return 0;
    if (a->L < 1 || a->L > VNEC_PATTERN_MAX) return (int)hipErrorInvalidValue;
    if (a->encoding < 0 || a->encoding > VNEC_ENC_F32_K) return (int)hipErrorInvalidValue;
    const int64_t blocks = (a->nslices + G_WAVES - 1) / G_WAVES;
    const int64_t grid = xcd_grid(blocks);
    if (a->wsize < 0 || a->wsize > VNEC_PATTERN_WIN_MAX || a->nwin < 0 || a->nwin > a->L)
        return (int)hipErrorInvalidValue;
    const bool win = a->wsize > 0;
    void (*k)(vnec_pattern_args, int64_t) =
        win ? (a->encoding == 0   ? k_spmv_pattern<8, 0, true>
               : a->encoding == 1 ? k_spmv_pattern<8, 1, true>
               : a->encoding == 2 ? k_spmv_pattern<8, 2, true>
                                  : k_spmv_pattern<8, VNEC_ENC_F32_K, true>)
            : (a->encoding == 0   ? k_spmv_pattern<8, 0, false>
               : a->encoding == 1 ? k_spmv_pattern<8, 1, false>
               : a->encoding == 2 ? k_spmv_pattern<8, 2, false>
                                  : k_spmv_pattern<8, VNEC_ENC_F32_K, false>);
    const size_t lds = win ? (a->encoding == VNEC_ENC_F32_K ? sizeof(float) : sizeof(double)) * (size_t)a->wsize : 0;
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(G_ROWS), lds, (hipStream_t)stream, *a, grid);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_scatter_rows(const double *ysub, const int32_t *rows, int64_t n, double *y,
                                     void *stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_scatter_rows<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, ysub,
                       rows, n, y);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_scatter_rows_f32(const float *ysub, const int32_t *rows, int64_t n, float *y,
                                         void *stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_scatter_rows<float>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, ysub,
                       rows, n, y);
    return (int)hipGetLastError();
}
