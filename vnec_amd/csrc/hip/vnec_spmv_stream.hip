// vnec_spmv_stream.hip — persistent, software-pipelined merge-path FP64 SpMV
// (variants L, E, C and H, one right-hand side), and the strided persistent
// form of the one-workgroup-per-tile kernel (k_spmv_strided, below).
//
// Each workgroup owns one contiguous run of merge-path tiles (workgroups on
// one XCD own neighbouring runs, so the x entries they share stay in that
// XCD's L2).  While a tile is reduced out of LDS, the loads of the next
// tile (values, columns, row ends, window runs) are already in flight in
// registers, so HBM never waits for the LDS phases.  The partial row at the
// end of a tile flows into the next tile of the same workgroup as the
// carry-in of its first thread; only the last partial row of each
// workgroup is left for the fix-up kernel, which adds them in workgroup
// order (deterministic; cf. spmv_vnec_L_for_mxv.c:160-164).
#include <hip/hip_runtime.h>

#include "vnec_common.h"
#include "vnec_hip.h"

// One nonzero's value as loaded: FP64, or FP32 hi (+ lo).  The FP32 forms
// are widened (and added) only where the product is formed: converting at
// the load would make the wavefront wait there for the next tile's loads,
// which the pipeline keeps in flight while the current tile is reduced.
template <int ENC> struct RawVal { double d; };
template <> struct RawVal<1> { float h, l; };
template <> struct RawVal<2> { float h; };
template <> struct RawVal<3> { float h; };  // the FP32 API: FP32 values, x and sums

template <int ENC>
__device__ __forceinline__ RawVal<ENC> ld_value(const vnec_merge_args &a, int64_t j);
template <> __device__ __forceinline__ RawVal<0> ld_value<0>(const vnec_merge_args &a, int64_t j)
{
    return {__builtin_nontemporal_load(a.val + j)};
}
template <> __device__ __forceinline__ RawVal<1> ld_value<1>(const vnec_merge_args &a, int64_t j)
{
    return {__builtin_nontemporal_load(a.hi + j), __builtin_nontemporal_load(a.lo + j)};
}
template <> __device__ __forceinline__ RawVal<2> ld_value<2>(const vnec_merge_args &a, int64_t j)
{
    return {__builtin_nontemporal_load(a.hi + j)};
}
template <> __device__ __forceinline__ RawVal<3> ld_value<3>(const vnec_merge_args &a, int64_t j)
{
    return {__builtin_nontemporal_load(a.hi + j)};
}
__device__ __forceinline__ double widen(RawVal<0> v) { return v.d; }
__device__ __forceinline__ double widen(RawVal<1> v) { return (double)v.h + (double)v.l; }
__device__ __forceinline__ double widen(RawVal<2> v) { return (double)v.h; }
__device__ __forceinline__ float widen(RawVal<3> v) { return v.h; }

// x, y and the per-tile carries in the kernel's value type (FP32 for the
// FP32 API, else FP64).
__device__ __forceinline__ const double *x_in(const vnec_merge_args &a, double) { return a.x; }
__device__ __forceinline__ const float *x_in(const vnec_merge_args &a, float) { return a.x32; }
__device__ __forceinline__ double *y_in(const vnec_merge_args &a, double) { return a.y; }
__device__ __forceinline__ float *y_in(const vnec_merge_args &a, float) { return a.y32; }
__device__ __forceinline__ double *carry_in(const vnec_merge_args &a, double) { return a.carry; }
__device__ __forceinline__ float *carry_in(const vnec_merge_args &a, float) { return a.carry32; }

// Registers holding one tile's streamed data (striped: item u * 256 + tid).
template <int IPT, bool WIN, int ENC = 0>
struct TileRegs {
    RawVal<ENC> v[IPT];
    int c[IPT];
    int re[IPT];
    int p[IPT];  // variant C: position of the nonzero in the tile's CSR range
    int cb;      // variant C: the tile's column base (packed relative columns)
};

// tid: the thread's index within its BS-thread tile pipeline.
template <int IPT, bool WIN, int ENC, bool SORT = false, int BS = VNEC_BLOCK>
__device__ __forceinline__ void load_tile(const vnec_merge_args &a, int64_t t, int2 c0, int2 c1,
                                          TileRegs<IPT, WIN, ENC> &R, int tid)
{
    const int nrow = c1.x - c0.x, nz = c1.y - c0.y;
    if (SORT) R.cb = a.cbase ? a.cbase[t] : 0;
#pragma unroll
    for (int u = 0; u < IPT; u++) {
        const int k = u * BS + tid;
        if (k < nz) {
            const int64_t j = (int64_t)c0.y + k;
            R.v[u] = ld_value<ENC>(a, j);
            R.c[u] = WIN ? (int)__builtin_nontemporal_load(a.lcol + j)
                         : __builtin_nontemporal_load(a.col + j);
            // (a position packed above the column is split where it is used,
            // so the loads of the next tile stay in flight)
            if (SORT && a.cpos_shift == 0) R.p[u] = (int)__builtin_nontemporal_load(a.spos + j);
        }
        if (k < nrow) R.re[u] = a.ptr[c0.x + 1 + k];
    }
}

// HOT > 0 (variant H): the a.nhot <= HOT most used columns of x, packed
// into a.xhot by k_gather_hot before the launch, are held in LDS for the
// whole launch; a.col then stores -1 - slot for those columns, so their
// gathers never leave the compute unit.
template <int HOT>
__device__ __forceinline__ double gather_x(const vnec_merge_args &a, const double *s_hot, int c)
{
    if (HOT > 0 && c < 0) return s_hot[-1 - c];
    return a.x[c];
}

// SORT (variant C): the tile's nonzeros arrive in column order, so each
// wavefront's x gathers are ascending and share cache lines where columns
// are dense; each product is written to its CSR position in LDS.
// BS: threads per tile (256, or VNEC_WIDE_BLOCK for the wide column-sorted
// tiles, whose larger column-order window shares more cache lines per
// gather).
template <int IPT, bool WIN, int ENC, int HOT = 0, bool SORT = false, int BS = VNEC_BLOCK>
__global__ __launch_bounds__(BS) void k_spmv_stream(vnec_merge_args a, int nwg,
                                                    int32_t *__restrict__ wg_row,
                                                    double *__restrict__ wg_carry)
{
    constexpr int T = BS * IPT;
    constexpr int NW = BS / VNEC_WAVE;
    __shared__ double s_prod[T];
    __shared__ int s_rowend[T];
    __shared__ double s_win[WIN ? T : 1];
    __shared__ double s_hot[HOT > 0 ? HOT : 1];
    __shared__ int s_f[NW + 1];
    __shared__ double s_v[NW + 1];
    __shared__ double s_carry;

    const int tid = threadIdx.x;
    const int rank = (int)xcd_tile(blockIdx.x, nwg);
    const int64_t tb = (int64_t)rank * a.ntiles / nwg;
    const int64_t te = (int64_t)(rank + 1) * a.ntiles / nwg;
    const int2 *co = (const int2 *)a.coords;

    if (tid == 0) s_carry = 0.0;
    if (tb >= te) {
        // An empty run reports the row at its boundary with a zero carry, so
        // the workgroups carrying into one row stay adjacent in rank order
        // (the fix-up relies on it to add each row's carries exactly once).
        if (tid == 0) {
            wg_row[rank] = co[tb].x;
            wg_carry[rank] = 0.0;
        }
        return;
    }

    // pipeline registers: current tile (A0, A1), next tile end (B1)
    // (plain if / else, not ?: — a select between a load and a register
    // would be lowered through a private-memory address)
    int2 A0 = co[tb], A1 = co[tb + 1];
    int2 B1 = A1;
    if (tb + 1 < te) B1 = co[tb + 2];
    int ra0 = 0, ra1 = 0, rb1 = 0, nec_a = 0, nec_b = 0;
    if (WIN) {
        ra0 = a.run_off[tb];
        ra1 = a.run_off[tb + 1];
        rb1 = ra1;
        nec_a = a.nec[tb];
        if (tb + 1 < te) {
            rb1 = a.run_off[tb + 2];
            nec_b = a.nec[tb + 1];
        }
    }
    TileRegs<IPT, WIN, ENC> cur, nxt;
    load_tile<IPT, WIN, ENC, SORT, BS>(a, tb, A0, A1, cur, tid);
    if (HOT > 0)
        for (int i = tid; i < a.nhot; i += BS) s_hot[i] = a.xhot[i];
    __syncthreads();

    for (int64_t t = tb; t < te; t++) {
        // --- issue the next tile's loads, and the coordinates after it
        int2 B2 = B1;
        int rb2 = rb1, nec_c = 0;
        if (t + 2 < te) {
            B2 = co[t + 3];
            if (WIN) {
                rb2 = a.run_off[t + 3];
                nec_c = a.nec[t + 2];
            }
        }
        if (t + 1 < te) load_tile<IPT, WIN, ENC, SORT, BS>(a, t + 1, A1, B1, nxt, tid);

        const int x0 = A0.x, y0 = A0.y;
        const int nrow = A1.x - A0.x, nz = A1.y - A0.y;

        // --- (E) compressed x window of this tile into LDS
        if (WIN) {
            fill_window(a.x, a.run_start, a.run_wpos, ra0, ra1, nec_a, s_win, (int *)s_prod);
            __syncthreads();
        }

        // --- row ends and products into LDS
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            const int k = u * BS + tid;
            if (k < nrow) s_rowend[k] = cur.re[u];
            if (k < nz) {
                int c = cur.c[u], pos = k;
                if (SORT) {
                    pos = cur.p[u];
                    if (a.cpos_shift > 0) {  // position packed above the column
                        pos = (int)((uint32_t)c >> a.cpos_shift);
                        c = (c & ((1 << a.cpos_shift) - 1)) + cur.cb;
                    }
                }
                s_prod[pos] = widen(cur.v[u]) * (WIN ? s_win[c] : gather_x<HOT>(a, s_hot, c));
            }
        }
        __syncthreads();

        // --- thread-level merge walk
        const int items = nrow + nz;
        const int d0 = tid * IPT < items ? tid * IPT : items;
        const int d1 = d0 + IPT < items ? d0 + IPT : items;
        int rx = (int)merge_search(s_rowend, nrow, y0, nz, d0);
        int ny = d0 - rx;
        double acc = 0.0, head_val = 0.0;
        int flag = 0, head_row = 0;
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            if (d0 + u < d1) {
                if (rx < nrow && s_rowend[rx] <= y0 + ny) {
                    if (!flag) {
                        head_row = x0 + rx;
                        head_val = acc;
                        flag = 1;
                    } else {
                        a.y[x0 + rx] = acc;
                    }
                    acc = 0.0;
                    rx++;
                } else {
                    acc += s_prod[ny];
                    ny++;
                }
            }
        }

        // --- segmented join across threads, carry-in from the previous tile
        int f = flag;
        double v = acc;
        wave_seg_scan_f64(f, v);
        const int w = wave_id();
        if (lane_id() == VNEC_WAVE - 1) {
            s_f[w] = f;
            s_v[w] = v;
        }
        __syncthreads();
        if (tid == 0) {
            double c = s_carry;
            for (int k = 0; k < NW; k++) {
                double tw = s_v[k];
                int fw = s_f[k];
                s_v[k] = c;
                c = fw ? tw : c + tw;
            }
            s_v[NW] = c;  // this tile's carry-out
        }
        __syncthreads();
        const double cw = s_v[w];
        const double blk = f ? v : v + cw;
        const double cin = wave_shift_up1_f64(blk, cw);
        if (flag) a.y[head_row] = head_val + cin;
        if (tid == 0) s_carry = s_v[NW];
        __syncthreads();

        // --- rotate the pipeline
        cur = nxt;
        A0 = A1;
        A1 = B1;
        B1 = B2;
        if (WIN) {
            ra0 = ra1;
            ra1 = rb1;
            rb1 = rb2;
            nec_a = nec_b;
            nec_b = nec_c;
        }
    }
    if (tid == 0) {
        wg_row[rank] = A0.x;  // end row of the last tile
        wg_carry[rank] = s_carry;
    }
}

// ------------------------------------------------ strided persistent form
// One-workgroup-per-tile semantics (every tile leaves its partial last row
// in carry[t] for k_carry_fixup; no carry flows between tiles) with a
// persistent, software-pipelined launch: the workgroups on XCD x sweep
// that XCD's contiguous share of the tiles together, workgroup j taking
// tiles j, j + W, j + 2W, ... of it (W workgroups per XCD), so the tiles in
// flight on one XCD are neighbours and share its L2, while each workgroup
// issues its next tile's loads before reducing the current one.
// BS: threads per tile (256, or VNEC_WIDE_BLOCK for wide column-sorted tiles).
// GB: each thread issues all of its x gathers of a tile before using any
// (more loads in flight per wavefront for more registers; the default waits
// for each gather in turn).
template <int IPT, int ENC, bool SORT, int BS = VNEC_BLOCK, bool GB = false>
__global__ __launch_bounds__(BS) void k_spmv_strided(vnec_merge_args a, int nwg)
{
    typedef typename enc_real<ENC>::type F;  // FP32 for the FP32 API's arithmetic
    constexpr int T = BS * IPT;
    constexpr int NW = BS / VNEC_WAVE;
    __shared__ F s_prod[T];
    __shared__ int s_rowend[T];
    __shared__ int s_f[NW + 1];
    __shared__ F s_v[NW + 1];
    const F *__restrict__ x = x_in(a, F());
    F *__restrict__ y = y_in(a, F());

    const int tid = threadIdx.x;
    const int64_t per = nwg / VNEC_XCDS;  // workgroups per XCD
    const int64_t xcd = blockIdx.x % VNEC_XCDS, j = blockIdx.x / VNEC_XCDS;
    const int64_t xlo = xcd * a.ntiles / VNEC_XCDS, xhi = (xcd + 1) * a.ntiles / VNEC_XCDS;
    const int2 *co = (const int2 *)a.coords;
    int64_t t = xlo + j;
    if (t >= xhi) return;

    int2 A0 = co[t], A1 = co[t + 1];
    TileRegs<IPT, false, ENC> cur, nxt;
    load_tile<IPT, false, ENC, SORT, BS>(a, t, A0, A1, cur, tid);

    for (; t < xhi; t += per) {
        // --- issue the next tile's loads
        const int64_t tn = t + per;
        int2 B0 = A0, B1 = A1;
        if (tn < xhi) {
            B0 = co[tn];
            B1 = co[tn + 1];
            load_tile<IPT, false, ENC, SORT, BS>(a, tn, B0, B1, nxt, tid);
        }

        const int x0 = A0.x, y0 = A0.y;
        const int nrow = A1.x - A0.x, nz = A1.y - A0.y;

        // --- row ends and products into LDS
        if (GB) {
            // the thread's gathers first (an item past the tile reads x[0],
            // present whenever the tile holds a nonzero, and is not used)
            F xv[IPT];
            int pv[IPT];
            if (nz > 0) {
#pragma unroll
                for (int u = 0; u < IPT; u++) {
                    const int k = u * BS + tid;
                    int c = cur.c[u], pos = k;
                    if (SORT) {
                        pos = cur.p[u];
                        if (a.cpos_shift > 0) {  // position packed above the column
                            pos = (int)((uint32_t)c >> a.cpos_shift);
                            c = (c & ((1 << a.cpos_shift) - 1)) + cur.cb;
                        }
                    }
                    pv[u] = pos;
                    xv[u] = x[k < nz ? c : 0];
                }
            }
#pragma unroll
            for (int u = 0; u < IPT; u++) {
                const int k = u * BS + tid;
                if (k < nrow) s_rowend[k] = cur.re[u];
                if (k < nz) s_prod[pv[u]] = widen(cur.v[u]) * xv[u];
            }
        } else {
#pragma unroll
            for (int u = 0; u < IPT; u++) {
                const int k = u * BS + tid;
                if (k < nrow) s_rowend[k] = cur.re[u];
                if (k < nz) {
                    int c = cur.c[u], pos = k;
                    if (SORT) {
                        pos = cur.p[u];
                        if (a.cpos_shift > 0) {  // position packed above the column
                            pos = (int)((uint32_t)c >> a.cpos_shift);
                            c = (c & ((1 << a.cpos_shift) - 1)) + cur.cb;
                        }
                    }
                    s_prod[pos] = widen(cur.v[u]) * x[c];
                }
            }
        }
        __syncthreads();

        // --- thread-level merge walk
        const int items = nrow + nz;
        const int d0 = tid * IPT < items ? tid * IPT : items;
        const int d1 = d0 + IPT < items ? d0 + IPT : items;
        int rx = (int)merge_search(s_rowend, nrow, y0, nz, d0);
        int ny = d0 - rx;
        F acc = 0, head_val = 0;
        int flag = 0, head_row = 0;
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            if (d0 + u < d1) {
                if (rx < nrow && s_rowend[rx] <= y0 + ny) {
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

        // --- segmented join across the tile's threads; the tile's carry-out
        F incl;
        const F cin = block_seg_carry<F, NW>(flag, acc, s_f, s_v, &incl);
        if (flag) y[head_row] = head_val + cin;
        if (tid == BS - 1) carry_in(a, F())[t] = incl;
        __syncthreads();

        cur = nxt;
        A0 = B0;
        A1 = B1;
    }
}

// Threads per workgroup of a launch: 256, or VNEC_WIDE_BLOCK for wide tiles.
static int block_threads(int block) { return block == VNEC_WIDE_BLOCK ? VNEC_WIDE_BLOCK : VNEC_BLOCK; }

typedef void (*strided_kernel_t)(vnec_merge_args, int);

static strided_kernel_t pick_strided(int variant, int ipt, int enc, int block, int gather)
{
    if (gather) {  // gathers issued together: column-sorted tiles, 3 to 7 items per thread
        if (variant != 6 || (enc != 0 && enc != 3)) return nullptr;
#define GCASE(I)                                                                                     \
    if (ipt == I && enc == 0)                                                                        \
        return block == VNEC_WIDE_BLOCK ? k_spmv_strided<I, 0, true, VNEC_WIDE_BLOCK, true>          \
                                        : k_spmv_strided<I, 0, true, VNEC_BLOCK, true>;              \
    if (ipt == I && enc == 3)                                                                        \
        return block == VNEC_WIDE_BLOCK ? k_spmv_strided<I, 3, true, VNEC_WIDE_BLOCK, true>          \
                                        : k_spmv_strided<I, 3, true, VNEC_BLOCK, true>;
        if (block != 0 && block != VNEC_BLOCK && block != VNEC_WIDE_BLOCK) return nullptr;
        GCASE(3) GCASE(5) GCASE(7)
#undef GCASE
        return nullptr;
    }
    if (block == VNEC_WIDE_BLOCK) {  // wide column-sorted tiles: C, FP64 or the FP32 API
        if (variant != 6 || (enc != 0 && enc != 3)) return nullptr;
#define WTCASE(I) \
    if (ipt == I) return enc == 3 ? k_spmv_strided<I, 3, true, VNEC_WIDE_BLOCK> : k_spmv_strided<I, 0, true, VNEC_WIDE_BLOCK>;
        WTCASE(3) WTCASE(5) WTCASE(7) WTCASE(9) WTCASE(11)
#undef WTCASE
        return nullptr;
    }
    if (block != 0 && block != VNEC_BLOCK) return nullptr;
#define TCASE(I, E)                                                           \
    if (ipt == I && enc == E) {                                               \
        if (variant == 0) return k_spmv_strided<I, E, false>;                 \
        if (variant == 6 && (E == 0 || E == 3)) return k_spmv_strided<I, E, true>; \
    }
#define TIPT(I) TCASE(I, 0) TCASE(I, 1) TCASE(I, 2) TCASE(I, 3)
    TIPT(3) TIPT(5) TIPT(7) TIPT(9) TIPT(11) TIPT(13) TIPT(15)
    return nullptr;
#undef TIPT
#undef TCASE
}

extern "C" int vnec_hip_strided_supported(int variant, int ipt, int encoding, int block, int gather)
{
    return pick_strided(variant, ipt, encoding, block, gather) != nullptr;
}

extern "C" int vnec_hip_strided_grid(int variant, int ipt, int encoding, int block, int gather, int *nwg)
{
    strided_kernel_t k = pick_strided(variant, ipt, encoding, block, gather);
    if (!k) return (int)hipErrorInvalidValue;
    int per_cu = 0, dev = 0, cus = 0;
    hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k, block_threads(block), 0);
    if (e != hipSuccess) return (int)e;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    int n = per_cu * cus;
    n = n / VNEC_XCDS * VNEC_XCDS;
    *nwg = n > 0 ? n : VNEC_XCDS;
    return 0;
}

extern "C" int vnec_hip_spmv_strided(const vnec_merge_args *a, int nwg, void *stream)
{
    strided_kernel_t k = pick_strided(a->variant, a->ipt, a->encoding, a->block, a->gather);
    if (!k || a->nrhs != 1 || nwg < VNEC_XCDS || nwg % VNEC_XCDS) return (int)hipErrorInvalidValue;
    if (a->ntiles <= 0) return 0;
    hipLaunchKernelGGL(k, dim3((unsigned)nwg), dim3((unsigned)block_threads(a->block)), 0, (hipStream_t)stream,
                       *a, nwg);
    return (int)hipGetLastError();
}

// ----------------------------------------------------------- variant H
// Four independent tile pipelines of 256 threads share one workgroup of
// 1024 threads and one LDS copy of the hot x entries, so the hot table
// costs no occupancy: per compute unit the same 16 wavefronts run as with
// the plain persistent kernel.  Pipeline g of workgroup b owns the run of
// tiles of rank 4 b' + g (b' the XCD-ordered workgroup rank); runs differ
// by at most one tile, and a pipeline whose run is over keeps meeting the
// workgroup barriers with empty tiles.
#define HOT_SUB 4

template <int IPT, int HOT>
__global__ __launch_bounds__(VNEC_BLOCK * HOT_SUB) void k_spmv_hot4(vnec_merge_args a, int nwg,
                                                                    int32_t *__restrict__ wg_row,
                                                                    double *__restrict__ wg_carry)
{
    constexpr int T = VNEC_BLOCK * IPT;
    __shared__ double s_hot[HOT];
    __shared__ double s_prod[HOT_SUB][T];
    __shared__ int s_rowend[HOT_SUB][T];
    __shared__ int s_f[HOT_SUB][VNEC_NWAVES + 1];
    __shared__ double s_v[HOT_SUB][VNEC_NWAVES + 1];
    __shared__ double s_carry[HOT_SUB];

    const int g = threadIdx.x / VNEC_BLOCK;
    const int tid = threadIdx.x % VNEC_BLOCK;
    const int w = tid / VNEC_WAVE;
    const int64_t nr = (int64_t)nwg * HOT_SUB;
    const int64_t base = xcd_tile(blockIdx.x, nwg) * HOT_SUB;
    const int64_t rank = base + g;
    const int64_t tb = rank * a.ntiles / nr, te = (rank + 1) * a.ntiles / nr;
    int64_t nmax = 0;
    for (int k = 0; k < HOT_SUB; k++) {
        const int64_t n = (base + k + 1) * a.ntiles / nr - (base + k) * a.ntiles / nr;
        nmax = n > nmax ? n : nmax;
    }
    const int2 *co = (const int2 *)a.coords;

    if (tid == 0) s_carry[g] = 0.0;
    for (int i = threadIdx.x; i < a.nhot; i += VNEC_BLOCK * HOT_SUB) s_hot[i] = a.xhot[i];

    int2 A0 = make_int2(0, 0), A1 = A0, B1 = A0;
    TileRegs<IPT, false> cur, nxt;
    if (tb < te) {
        A0 = co[tb];
        A1 = co[tb + 1];
        B1 = A1;
        if (tb + 1 < te) B1 = co[tb + 2];
        load_tile<IPT, false, 0>(a, tb, A0, A1, cur, tid);
    }
    __syncthreads();

    for (int64_t step = 0; step < nmax; step++) {
        const int64_t t = tb + step;
        const bool act = t < te;
        int2 B2 = B1;
        if (act && t + 2 < te) B2 = co[t + 3];
        if (act && t + 1 < te) load_tile<IPT, false, 0>(a, t + 1, A1, B1, nxt, tid);

        const int x0 = A0.x, y0 = A0.y;
        const int nrow = act ? A1.x - A0.x : 0, nz = act ? A1.y - A0.y : 0;
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            const int k = u * VNEC_BLOCK + tid;
            if (k < nrow) s_rowend[g][k] = cur.re[u];
            if (k < nz) {
                const int c = cur.c[u];
                s_prod[g][k] = widen(cur.v[u]) * (c < 0 ? s_hot[-1 - c] : a.x[c]);
            }
        }
        __syncthreads();

        const int items = nrow + nz;
        const int d0 = tid * IPT < items ? tid * IPT : items;
        const int d1 = d0 + IPT < items ? d0 + IPT : items;
        int rx = (int)merge_search(s_rowend[g], nrow, y0, nz, d0);
        int ny = d0 - rx;
        double acc = 0.0, head_val = 0.0;
        int flag = 0, head_row = 0;
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            if (d0 + u < d1) {
                if (rx < nrow && s_rowend[g][rx] <= y0 + ny) {
                    if (!flag) {
                        head_row = x0 + rx;
                        head_val = acc;
                        flag = 1;
                    } else {
                        a.y[x0 + rx] = acc;
                    }
                    acc = 0.0;
                    rx++;
                } else {
                    acc += s_prod[g][ny];
                    ny++;
                }
            }
        }

        int f = flag;
        double v = acc;
        wave_seg_scan_f64(f, v);
        if (lane_id() == VNEC_WAVE - 1) {
            s_f[g][w] = f;
            s_v[g][w] = v;
        }
        __syncthreads();
        if (tid == 0) {
            double c = s_carry[g];
            for (int k = 0; k < VNEC_NWAVES; k++) {
                const double tw = s_v[g][k];
                const int fw = s_f[g][k];
                s_v[g][k] = c;
                c = fw ? tw : c + tw;
            }
            s_v[g][VNEC_NWAVES] = c;
        }
        __syncthreads();
        const double cw = s_v[g][w];
        const double blk = f ? v : v + cw;
        const double cin = wave_shift_up1_f64(blk, cw);
        if (flag) a.y[head_row] = head_val + cin;
        if (tid == 0 && act) s_carry[g] = s_v[g][VNEC_NWAVES];
        __syncthreads();

        if (act) {
            cur = nxt;
            A0 = A1;
            A1 = B1;
            B1 = B2;
        }
    }
    if (tid == 0) {
        if (tb < te) {
            wg_row[rank] = A0.x;
            wg_carry[rank] = s_carry[g];
        } else {
            wg_row[rank] = co[tb].x;
            wg_carry[rank] = 0.0;
        }
    }
}

// ----------------------------------------------------------- variant E
// Pipelined form of the compressed-window kernel.  In the iteration that
// reduces tile t, the workgroup also
//   - loads the run descriptors of tile t + 2 (one run per thread),
//   - loads the values / local columns / row ends of tile t + 1,
//   - gathers tile t + 1's x window into registers (each thread finds the
//     run of its window positions in the descriptors staged in LDS one
//     iteration earlier),
// so that the only work on a tile's critical path is in LDS.  A tile may
// hold at most VNEC_ERUNS runs (the plan falls back to the one-workgroup-
// per-tile kernel otherwise).
#define VNEC_ERUNS VNEC_BLOCK
static_assert(VNEC_ERUNS == VNEC_E_STREAM_MAX_RUNS, "one run descriptor per thread");

__device__ __forceinline__ int find_run(const int2 *runs, int nr, int p)
{
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (runs[mid].y <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int IPT, int ENC>
struct ERegs {
    RawVal<ENC> v[IPT];
    double xw[IPT];
    int c[IPT];
    int re[IPT];
};

template <int IPT, int ENC>
__device__ __forceinline__ void e_load(const vnec_merge_args &a, int2 c0, int2 c1, int nec,
                                       const int2 *runs, int nr, ERegs<IPT, ENC> &R)
{
    const int tid = threadIdx.x;
    const int nrow = c1.x - c0.x, nz = c1.y - c0.y;
#pragma unroll
    for (int u = 0; u < IPT; u++) {
        const int k = u * VNEC_BLOCK + tid;
        if (k < nz) {
            const int64_t j = (int64_t)c0.y + k;
            R.v[u] = ld_value<ENC>(a, j);
            R.c[u] = (int)__builtin_nontemporal_load(a.lcol + j);
        }
        if (k < nrow) R.re[u] = a.ptr[c0.x + 1 + k];
        if (k < nec) {
            const int2 r = runs[find_run(runs, nr, k)];
            R.xw[u] = a.x[r.x + (k - r.y)];
        }
    }
}

template <int IPT, int ENC>
__global__ __launch_bounds__(VNEC_BLOCK) void k_spmv_estream(vnec_merge_args a, int nwg,
                                                             int32_t *__restrict__ wg_row,
                                                             double *__restrict__ wg_carry)
{
    constexpr int T = VNEC_BLOCK * IPT;
    __shared__ double s_prod[T];
    __shared__ int s_rowend[T];
    __shared__ double s_win[T];
    __shared__ int2 s_runs[2][VNEC_ERUNS];
    __shared__ int s_f[VNEC_NWAVES + 1];
    __shared__ double s_v[VNEC_NWAVES + 1];
    __shared__ double s_carry;

    const int tid = threadIdx.x;
    const int rank = (int)xcd_tile(blockIdx.x, nwg);
    const int64_t tb = (int64_t)rank * a.ntiles / nwg;
    const int64_t te = (int64_t)(rank + 1) * a.ntiles / nwg;
    const int2 *co = (const int2 *)a.coords;

    if (tid == 0) s_carry = 0.0;
    if (tb >= te) {
        if (tid == 0) {
            wg_row[rank] = co[tb].x;
            wg_carry[rank] = 0.0;
        }
        return;
    }

    // scalars of tiles t .. t + 3 (if / else only: see the L kernel)
    int2 cA = co[tb], cB = co[tb + 1], cC = cB;
    int roA = a.run_off[tb], roB = a.run_off[tb + 1], roC = roB, roD = roB;
    int necA = a.nec[tb], necB = 0;
    if (tb + 1 < te) {
        cC = co[tb + 2];
        roC = a.run_off[tb + 2];
        roD = roC;
        necB = a.nec[tb + 1];
    }
    if (tb + 2 < te) roD = a.run_off[tb + 3];
    // run descriptors of tiles tb and tb + 1
    if (tid < roB - roA) s_runs[tb & 1][tid] = make_int2(a.run_start[roA + tid], a.run_wpos[roA + tid]);
    if (tb + 1 < te && tid < roC - roB)
        s_runs[(tb + 1) & 1][tid] = make_int2(a.run_start[roB + tid], a.run_wpos[roB + tid]);
    __syncthreads();
    ERegs<IPT, ENC> cur, nxt;
    e_load<IPT, ENC>(a, cA, cB, necA, s_runs[tb & 1], roB - roA, cur);

    for (int64_t t = tb; t < te; t++) {
        // scalars for the next iteration
        int2 cD = cC;
        int roE = roD, necC = 0;
        if (t + 2 < te) {
            cD = co[t + 3];
            necC = a.nec[t + 2];
        }
        if (t + 3 < te) roE = a.run_off[t + 4];
        // run descriptors of tile t + 2 (registers until the end of the iteration)
        int2 run2 = make_int2(0, 0);
        const bool has2 = t + 2 < te && tid < roD - roC;
        if (has2) run2 = make_int2(a.run_start[roC + tid], a.run_wpos[roC + tid]);
        // tile t + 1: streamed data and its x window
        if (t + 1 < te) e_load<IPT, ENC>(a, cB, cC, necB, s_runs[(t + 1) & 1], roC - roB, nxt);

        // --- reduce tile t
        const int x0 = cA.x, y0 = cA.y;
        const int nrow = cB.x - cA.x, nz = cB.y - cA.y;
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            const int k = u * VNEC_BLOCK + tid;
            if (k < necA) s_win[k] = cur.xw[u];
            if (k < nrow) s_rowend[k] = cur.re[u];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            const int k = u * VNEC_BLOCK + tid;
            if (k < nz) s_prod[k] = widen(cur.v[u]) * s_win[cur.c[u]];
        }
        __syncthreads();
        const int items = nrow + nz;
        const int d0 = tid * IPT < items ? tid * IPT : items;
        const int d1 = d0 + IPT < items ? d0 + IPT : items;
        int rx = (int)merge_search(s_rowend, nrow, y0, nz, d0);
        int ny = d0 - rx;
        double acc = 0.0, head_val = 0.0;
        int flag = 0, head_row = 0;
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            if (d0 + u < d1) {
                if (rx < nrow && s_rowend[rx] <= y0 + ny) {
                    if (!flag) {
                        head_row = x0 + rx;
                        head_val = acc;
                        flag = 1;
                    } else {
                        a.y[x0 + rx] = acc;
                    }
                    acc = 0.0;
                    rx++;
                } else {
                    acc += s_prod[ny];
                    ny++;
                }
            }
        }
        int f = flag;
        double v = acc;
        wave_seg_scan_f64(f, v);
        const int w = wave_id();
        if (lane_id() == VNEC_WAVE - 1) {
            s_f[w] = f;
            s_v[w] = v;
        }
        __syncthreads();
        if (tid == 0) {
            double c = s_carry;
            for (int k = 0; k < VNEC_NWAVES; k++) {
                double tw = s_v[k];
                int fw = s_f[k];
                s_v[k] = c;
                c = fw ? tw : c + tw;
            }
            s_v[VNEC_NWAVES] = c;
        }
        __syncthreads();
        const double cw = s_v[w];
        const double blk = f ? v : v + cw;
        const double cin = wave_shift_up1_f64(blk, cw);
        if (flag) a.y[head_row] = head_val + cin;
        if (tid == 0) s_carry = s_v[VNEC_NWAVES];
        // tile t + 2's runs go to the buffer tile t used
        if (has2) s_runs[t & 1][tid] = run2;
        __syncthreads();

        cur = nxt;
        cA = cB;
        cB = cC;
        cC = cD;
        roA = roB;
        roB = roC;
        roC = roD;
        roD = roE;
        necA = necB;
        necB = necC;
    }
    if (tid == 0) {
        wg_row[rank] = cA.x;
        wg_carry[rank] = s_carry;
    }
}

// Adds each workgroup's carry to its row: the first workgroup of each run
// carrying into one row sums the run in order and adds it once.
__global__ void k_wg_fixup(const int32_t *__restrict__ wg_row, const double *__restrict__ wg_carry,
                           int nwg, double *__restrict__ y, int64_t nrows)
{
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwg) return;
    int row = wg_row[w];
    if (row >= nrows) return;
    if (w > 0 && wg_row[w - 1] == row) return;
    // the run's first four candidates are loaded together, then summed in
    // workgroup order; longer runs continue one by one
    int rr[4];
    double cc[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const bool ok = w + i < nwg;
        rr[i] = i == 0 ? row : (ok ? wg_row[w + i] : -1);
        cc[i] = ok ? wg_carry[w + i] : 0.0;
    }
    double s = 0.0;
    bool in = true;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        in = in && rr[i] == row;
        if (in) s += cc[i];
    }
    if (in)
        for (int u = w + 4; u < nwg && wg_row[u] == row; u++) s += wg_carry[u];
    y[row] += s;
}

typedef void (*stream_kernel_t)(vnec_merge_args, int, int32_t *, double *);

static stream_kernel_t pick_stream(int variant, int ipt, int enc, int block)
{
    if (block == VNEC_WIDE_BLOCK) {  // wide column-sorted tiles: FP64, C
        if (variant != 6 || enc != 0) return nullptr;
#define WCASE(I) \
    if (ipt == I) return k_spmv_stream<I, false, 0, 0, true, VNEC_WIDE_BLOCK>;
        WCASE(3) WCASE(5) WCASE(7) WCASE(9) WCASE(11)
#undef WCASE
        return nullptr;
    }
    if (block != 0 && block != VNEC_BLOCK) return nullptr;
#define SCASE(I, E)                                                  \
    if (ipt == I && enc == E) {                                      \
        if (variant == 0) return k_spmv_stream<I, false, E>;         \
        if (variant == 1) return k_spmv_estream<I, E>;               \
        if (variant == 6 && E == 0) return k_spmv_stream<I, false, 0, 0, true>; \
    }
#define SIPT(I) SCASE(I, 0) SCASE(I, 1) SCASE(I, 2)
    SIPT(3) SIPT(5) SIPT(7) SIPT(9) SIPT(11) SIPT(13) SIPT(15)
    return nullptr;
#undef SIPT
#undef SCASE
}


extern "C" int vnec_hip_stream_supported(int variant, int ipt, int encoding, int block)
{
    return (variant == 0 || variant == 1 || variant == 6) && pick_stream(variant, ipt, encoding, block) != nullptr;
}

extern "C" int vnec_hip_stream_grid(int variant, int ipt, int encoding, int block, int *nwg)
{
    stream_kernel_t k = (variant == 0 || variant == 1 || variant == 6) ? pick_stream(variant, ipt, encoding, block)
                                                                        : nullptr;
    if (!k) return (int)hipErrorInvalidValue;
    int per_cu = 0, dev = 0, cus = 0;
    hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k, block_threads(block), 0);
    if (e != hipSuccess) return (int)e;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    int n = per_cu * cus;
    n = n / VNEC_XCDS * VNEC_XCDS;
    *nwg = n > 0 ? n : VNEC_XCDS;
    return 0;
}

extern "C" int vnec_hip_spmv_stream(const vnec_merge_args *a, int nwg, int32_t *wg_row,
                                    double *wg_carry, void *stream)
{
    stream_kernel_t k = pick_stream(a->variant, a->ipt, a->encoding, a->block);
    if (!k || a->nrhs != 1 || nwg < VNEC_XCDS || nwg % VNEC_XCDS) return (int)hipErrorInvalidValue;
    if (a->ntiles <= 0) return 0;
    hipLaunchKernelGGL(k, dim3((unsigned)nwg), dim3((unsigned)block_threads(a->block)), 0, (hipStream_t)stream,
                       *a, nwg, wg_row, wg_carry);
    return (int)hipGetLastError();
}

// ----------------------------------------------------------- variant H
static stream_kernel_t pick_hot(int ipt, int hot)
{
#define HCASE(I, H) \
    if (ipt == I && hot == H) return k_spmv_hot4<I, H>;
    HCASE(5, 4096) HCASE(7, 4096) HCASE(5, 8192) HCASE(7, 8192)
#undef HCASE
    return nullptr;
}

extern "C" int vnec_hip_hot_supported(int ipt, int hot) { return pick_hot(ipt, hot) != nullptr; }

extern "C" int vnec_hip_hot_grid(int ipt, int hot, int *nwg)
{
    stream_kernel_t k = pick_hot(ipt, hot);
    if (!k) return (int)hipErrorInvalidValue;
    int per_cu = 0, dev = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k,
                                                                VNEC_BLOCK * HOT_SUB, 0);
    if (e != hipSuccess) return (int)e;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    int n = per_cu * cus;
    n = n / VNEC_XCDS * VNEC_XCDS;
    *nwg = n > 0 ? n : VNEC_XCDS;
    return 0;
}

extern "C" int vnec_hip_spmv_hot(const vnec_merge_args *a, int hot, int nwg, int32_t *wg_row,
                                 double *wg_carry, void *stream)
{
    stream_kernel_t k = pick_hot(a->ipt, hot);
    if (!k || a->nrhs != 1 || a->encoding != 0 || a->nhot > hot || nwg < VNEC_XCDS ||
        nwg % VNEC_XCDS)
        return (int)hipErrorInvalidValue;
    if (a->ntiles <= 0) return 0;
    hipLaunchKernelGGL(k, dim3((unsigned)nwg), dim3(VNEC_BLOCK * HOT_SUB), 0, (hipStream_t)stream, *a,
                       nwg, wg_row, wg_carry);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_wg_fixup(const int32_t *wg_row, const double *wg_carry, int nwg, double *y,
                                 int64_t nrows, void *stream)
{
    if (nwg <= 0) return 0;
    hipLaunchKernelGGL(k_wg_fixup, dim3((unsigned)((nwg + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, wg_row, wg_carry, nwg, y, nrows);
    return (int)hipGetLastError();
}
