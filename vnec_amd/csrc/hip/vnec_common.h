// vnec_common.h — device helpers shared by the VNEC HIP kernels (gfx950,
// wave64): DPP wavefront scans, block scans, XCD-aware tile mapping and the
// merge-path search.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define VNEC_WAVE 64
#define VNEC_BLOCK 256
#define VNEC_NWAVES (VNEC_BLOCK / VNEC_WAVE)
#define VNEC_XCDS 8

// DPP controls (GFX9 encoding): row_shr:n = 0x110 + n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143.
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

__device__ __forceinline__ int lane_id() { return threadIdx.x & (VNEC_WAVE - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / VNEC_WAVE; }

// Lane l receives src from lane l - shift inside its 16-lane row (rows via
// bcast), or 0 where there is no source: the identity of every scan below.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_i32(int src)
{
    return __builtin_amdgcn_update_dpp(0, src, CTRL, ROW_MASK, 0xF, true);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double src)
{
    union { double d; int i[2]; } u, r;
    u.d = src;
    r.i[0] = __builtin_amdgcn_update_dpp(0, u.i[0], CTRL, ROW_MASK, 0xF, true);
    r.i[1] = __builtin_amdgcn_update_dpp(0, u.i[1], CTRL, ROW_MASK, 0xF, true);
    return r.d;
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f32(float src)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(src), CTRL, ROW_MASK, 0xF, true));
}

// One DPP move of a value of either real type.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_real(double v) { return dpp_f64<CTRL, ROW_MASK>(v); }
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_real(float v) { return dpp_f32<CTRL, ROW_MASK>(v); }

// Inclusive wavefront sum scan of an int.
__device__ __forceinline__ int wave_incl_scan_i32(int v)
{
    v += dpp_i32<DPP_ROW_SHR(1), 0xF>(v);
    v += dpp_i32<DPP_ROW_SHR(2), 0xF>(v);
    v += dpp_i32<DPP_ROW_SHR(4), 0xF>(v);
    v += dpp_i32<DPP_ROW_SHR(8), 0xF>(v);
    v += dpp_i32<DPP_ROW_BCAST15, 0xA>(v);
    v += dpp_i32<DPP_ROW_BCAST31, 0xC>(v);
    return v;
}

// Inclusive segmented wavefront scan: the pair (f, v) combines as
// (f1, v1) . (f2, v2) = (f1 | f2, f2 ? v2 : v1 + v2), i.e. a set flag
// restarts the running sum at that lane's own value.
template <int CTRL, int ROW_MASK, typename F>
__device__ __forceinline__ void seg_step(int &f, F &v)
{
    int fo = dpp_i32<CTRL, ROW_MASK>(f);
    F vo = dpp_real<CTRL, ROW_MASK>(v);
    if (!f) v += vo;
    f |= fo;
}

// F = double or float.
template <typename F>
__device__ __forceinline__ void wave_seg_scan(int &f, F &v)
{
    seg_step<DPP_ROW_SHR(1), 0xF>(f, v);
    seg_step<DPP_ROW_SHR(2), 0xF>(f, v);
    seg_step<DPP_ROW_SHR(4), 0xF>(f, v);
    seg_step<DPP_ROW_SHR(8), 0xF>(f, v);
    seg_step<DPP_ROW_BCAST15, 0xA>(f, v);
    seg_step<DPP_ROW_BCAST31, 0xC>(f, v);
}

__device__ __forceinline__ void wave_seg_scan_f64(int &f, double &v) { wave_seg_scan(f, v); }

// Value of lane (l - 1) (lane 0 gets `first`).
__device__ __forceinline__ double wave_shift_up1_f64(double v, double first)
{
    union { double d; int i[2]; } u, r;
    u.d = v;
    r.i[0] = __shfl_up(u.i[0], 1, VNEC_WAVE);
    r.i[1] = __shfl_up(u.i[1], 1, VNEC_WAVE);
    return lane_id() == 0 ? first : r.d;
}

__device__ __forceinline__ float wave_shift_up1_f32(float v, float first)
{
    float r = __shfl_up(v, 1, VNEC_WAVE);
    return lane_id() == 0 ? first : r;
}

__device__ __forceinline__ double wave_shift_up1(double v, double first) { return wave_shift_up1_f64(v, first); }
__device__ __forceinline__ float wave_shift_up1(float v, float first) { return wave_shift_up1_f32(v, first); }

__device__ __forceinline__ int wave_shift_up1_i32(int v, int first)
{
    int r = __shfl_up(v, 1, VNEC_WAVE);
    return lane_id() == 0 ? first : r;
}

// Block-wide exclusive sum scan of one int per thread; *total receives
// the block sum.  s_tmp holds VNEC_NWAVES + 1 ints.
__device__ __forceinline__ int block_excl_scan_i32(int v, int *s_tmp, int *total)
{
    int inc = wave_incl_scan_i32(v);
    if (lane_id() == VNEC_WAVE - 1) s_tmp[wave_id()] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < VNEC_NWAVES; w++) {
            int t = s_tmp[w];
            s_tmp[w] = run;
            run += t;
        }
        s_tmp[VNEC_NWAVES] = run;
    }
    __syncthreads();
    int r = inc - v + s_tmp[wave_id()];
    *total = s_tmp[VNEC_NWAVES];
    __syncthreads();
    return r;
}

// Blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8).
// Map them so that each XCD walks one contiguous run of tiles: neighbouring
// tiles share x entries, which then stay in that XCD's L2.  The grid must
// be a multiple of 8; tiles >= ntiles are idle.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t grid)
{
    int64_t chunk = grid / VNEC_XCDS;
    return (b % VNEC_XCDS) * chunk + b / VNEC_XCDS;
}

static inline int64_t xcd_grid(int64_t ntiles)
{
    return (ntiles + VNEC_XCDS - 1) / VNEC_XCDS * VNEC_XCDS;
}

// Gathers a tile's compressed x window (its distinct columns, stored as
// runs of consecutive columns: run r starts at column run_start[r] and
// window position run_wpos[r]) into s_win[0 .. nec).
// Long runs: wavefront w copies runs w, w + 4, ..., lanes reading
// consecutive x entries (coalesced; run descriptors are wave-uniform, read
// with scalar loads).  Short runs (the window of a scattered tile): one
// lane per window position, finding its run by binary search over the run
// descriptors staged in s_runs (2 * nr ints of LDS; with s_runs == nullptr
// the long-run form is always used).
template <typename F>
__device__ __forceinline__ void fill_window(const F *__restrict__ x,
                                            const int32_t *__restrict__ run_start,
                                            const int32_t *__restrict__ run_wpos, int r0, int r1,
                                            int nec, F *s_win, int *s_runs)
{
    const int nr = r1 - r0;
    if (nr <= 0) return;
    if (nec >= 16 * nr || s_runs == nullptr) {
        const int w = __builtin_amdgcn_readfirstlane(wave_id());
        for (int r = r0 + w; r < r1; r += VNEC_NWAVES) {
            const int start = run_start[r], wpos = run_wpos[r];
            int end = nec;
            if (r + 1 < r1) end = run_wpos[r + 1];
            for (int k = lane_id(); k < end - wpos; k += VNEC_WAVE) s_win[wpos + k] = x[start + k];
        }
    } else {
        int *s_rs = s_runs, *s_rw = s_runs + nr;
        for (int i = threadIdx.x; i < nr; i += VNEC_BLOCK) {
            s_rs[i] = run_start[r0 + i];
            s_rw[i] = run_wpos[r0 + i];
        }
        __syncthreads();
        for (int p = threadIdx.x; p < nec; p += VNEC_BLOCK) {
            int lo = 0, hi = nr - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (s_rw[mid] <= p) lo = mid; else hi = mid - 1;
            }
            s_win[p] = x[s_rs[lo] + (p - s_rw[lo])];
        }
    }
}

// Joins per-thread (flag, tail) pairs of a segmented sum across the
// workgroup.  Returns the exclusive carry into this thread; *incl (if not
// NULL) gets the inclusive value (the last thread's closes the tile).
// s_f / s_v: NW + 1 entries (NW: wavefronts in the workgroup).
template <typename F, int NW = VNEC_NWAVES>
__device__ __forceinline__ F block_seg_carry(int flag, F tail, int *s_f, F *s_v, F *incl)
{
    int f = flag;
    F v = tail;
    wave_seg_scan(f, v);
    const int w = wave_id();
    if (lane_id() == VNEC_WAVE - 1) {
        s_f[w] = f;
        s_v[w] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        F c = 0;
        for (int k = 0; k < NW; k++) {
            const F tw = s_v[k];
            const int fw = s_f[k];
            s_v[k] = c;  // carry into wave k
            c = fw ? tw : c + tw;
        }
    }
    __syncthreads();
    const F cw = s_v[w];
    const F blk = f ? v : v + cw;
    if (incl) *incl = blk;
    return wave_shift_up1(blk, cw);
}

// Value type a kernel computes in for a value encoding: encodings 0-2
// (FP64 values, FP32 hi + lo, FP32 hi) compute in FP64; encoding 3 is the
// FP32 arithmetic of the FP32 API (FP32 values, x, y and sums).
#define VNEC_ENC_F32_K 3
template <int ENC> struct enc_real { typedef double type; };
template <> struct enc_real<VNEC_ENC_F32_K> { typedef float type; };

// Merge-path search (spmv_vnec_common.c:3-33 semantics): merge list A =
// row ends re[0..na), list B = b0 + 0..nb; returns the number of A items
// consumed on diagonal d, consuming A first on ties (A[i] <= B[j]).
template <typename RowEnd>
__device__ __forceinline__ int64_t merge_search(RowEnd re, int64_t na, int64_t b0, int64_t nb,
                                                int64_t d)
{
    int64_t lo = d - nb > 0 ? d - nb : 0;
    int64_t hi = d < na ? d : na;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((int64_t)re[mid] <= b0 + d - mid - 1)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
