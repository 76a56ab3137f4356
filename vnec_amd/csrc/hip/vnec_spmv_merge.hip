// vnec_spmv_merge.hip — merge-path SpMV for gfx950, one workgroup per tile
// (variants L, E and C; multi-RHS; the FP32 API's L) and the carry fix-up.
//
// Restates the merge-based CSR of spmv_vnec_L_for_mxv.c:124-165 for wave64
// workgroups: a tile is a fixed share of the merge list (row ends +
// nonzeros) found by the GPU partitioner, so every workgroup does the same
// amount of work whatever the row lengths.  Per tile:
//   1. (E only) the tile's compressed x window — its distinct columns,
//      described as runs of consecutive columns — is gathered into LDS
//      (the GPU form of the compact-x gather, spmv_vnec_D_for_mxv.c:416-436);
//   2. row ends and the products a_j * x(col_j) are staged in LDS with
//      coalesced loads (values FP64, or FP32 hi + lo / hi only, all summed
//      in FP64; or, for the FP32 API, FP32 values, x and sums);
//   3. each thread walks IPT merge items sequentially, writing rows it
//      completes and keeping the partial sums at its two ends;
//   4. a segmented scan over the threads (DPP within the wavefront, LDS
//      across the four wavefronts) joins rows that span threads;
//   5. the tile's partial last row is left in carry[t]; a second kernel
//      adds the carries to y in tile order (deterministic fix-up).
#include <hip/hip_runtime.h>
#include <climits>

#include "vnec_common.h"
#include "vnec_hip.h"

template <int ENC>
__device__ __forceinline__ typename enc_real<ENC>::type load_value(const vnec_merge_args &a, int64_t j)
{
    if (ENC == 0) return __builtin_nontemporal_load(a.val + j);
    if (ENC == 1)
        return (double)__builtin_nontemporal_load(a.hi + j) +
               (double)__builtin_nontemporal_load(a.lo + j);
    return __builtin_nontemporal_load(a.hi + j);  // FP32 hi (FP64 sums), or FP32 values
}

// x, y and the tile carries in the kernel's value type.
__device__ __forceinline__ const double *x_of(const vnec_merge_args &a, double) { return a.x; }
__device__ __forceinline__ const float *x_of(const vnec_merge_args &a, float) { return a.x32; }
__device__ __forceinline__ double *y_of(const vnec_merge_args &a, double) { return a.y; }
__device__ __forceinline__ float *y_of(const vnec_merge_args &a, float) { return a.y32; }
__device__ __forceinline__ double *carry_of(const vnec_merge_args &a, double) { return a.carry; }
__device__ __forceinline__ float *carry_of(const vnec_merge_args &a, float) { return a.carry32; }

// SORT (variant C): the tile's nonzeros arrive in column order with their
// CSR offsets in a.spos; each product is staged at its CSR offset.
template <int IPT, bool WIN, int ENC, bool SORT = false>
__global__ __launch_bounds__(VNEC_BLOCK) void k_spmv_merge1(vnec_merge_args a, int64_t grid)
{
    typedef typename enc_real<ENC>::type F;
    constexpr int T = VNEC_BLOCK * IPT;
    __shared__ F s_prod[T];
    __shared__ int s_rowend[T];
    __shared__ F s_win[WIN ? T : 1];
    __shared__ int s_f[VNEC_NWAVES + 1];
    __shared__ F s_v[VNEC_NWAVES + 1];

    const int64_t t = xcd_tile(blockIdx.x, grid);
    if (t >= a.ntiles) return;
    const F *__restrict__ x = x_of(a, F());
    F *__restrict__ y = y_of(a, F());
    const int2 *co = (const int2 *)a.coords;
    const int2 c0 = co[t], c1 = co[t + 1];
    const int cbt = (SORT && a.cbase) ? a.cbase[t] : 0;  // column base of packed relative columns
    const int x0 = c0.x, y0 = c0.y;
    const int nrow = c1.x - c0.x, nz = c1.y - c0.y;
    const int tid = threadIdx.x;

    if (WIN) {
        // 1. window fill (run descriptors staged in s_prod when needed)
        fill_window(x, a.run_start, a.run_wpos, a.run_off[t], a.run_off[t + 1], a.nec[t], s_win,
                    (int *)s_prod);
        __syncthreads();
    }

    // 2. row ends and products
    for (int i = tid; i < nrow; i += VNEC_BLOCK) s_rowend[i] = a.ptr[x0 + 1 + i];
    {
        F av[IPT];
        int cv[IPT], pv[IPT];
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            int k = u * VNEC_BLOCK + tid;
            pv[u] = k;
            if (k < nz) {
                int64_t j = (int64_t)y0 + k;
                av[u] = load_value<ENC>(a, j);
                cv[u] = WIN ? (int)__builtin_nontemporal_load(a.lcol + j)
                            : __builtin_nontemporal_load(a.col + j);
                if (SORT && a.cpos_shift == 0) pv[u] = (int)__builtin_nontemporal_load(a.spos + j);
            }
        }
#pragma unroll
        for (int u = 0; u < IPT; u++) {
            int k = u * VNEC_BLOCK + tid;
            if (k < nz) {
                int c = cv[u], pos = pv[u];
                if (SORT && a.cpos_shift > 0) {  // position packed above the column
                    pos = (int)((uint32_t)c >> a.cpos_shift);
                    c = (c & ((1 << a.cpos_shift) - 1)) + cbt;
                }
                s_prod[pos] = av[u] * (WIN ? s_win[c] : x[c]);
            }
        }
    }
    __syncthreads();

    // 3. thread-level merge walk over IPT items
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

    // 4. join rows that span threads
    F incl;
    F cin = block_seg_carry(flag, acc, s_f, s_v, &incl);
    if (flag) y[head_row] = head_val + cin;
    // 5. tile carry (partial row c1.x)
    if (tid == VNEC_BLOCK - 1) carry_of(a, F())[t] = incl;
}

// Multi-RHS (variant L): x is [ncols][K], y is [nrows][K].  Values and
// columns are staged in LDS; the walk gathers K contiguous x entries.
template <int IPT, int K, int ENC>
__global__ __launch_bounds__(VNEC_BLOCK) void k_spmv_mergeK(vnec_merge_args a, int64_t grid)
{
    constexpr int T = VNEC_BLOCK * IPT;
    __shared__ double s_val[T];
    __shared__ int s_col[T];
    __shared__ int s_rowend[T];
    __shared__ int s_f[VNEC_NWAVES + 1];
    __shared__ double s_v[VNEC_NWAVES + 1];

    const int64_t t = xcd_tile(blockIdx.x, grid);
    if (t >= a.ntiles) return;
    const int2 *co = (const int2 *)a.coords;
    const int2 c0 = co[t], c1 = co[t + 1];
    const int x0 = c0.x, y0 = c0.y;
    const int nrow = c1.x - c0.x, nz = c1.y - c0.y;
    const int tid = threadIdx.x;

    for (int i = tid; i < nrow; i += VNEC_BLOCK) s_rowend[i] = a.ptr[x0 + 1 + i];
#pragma unroll
    for (int u = 0; u < IPT; u++) {
        int k = u * VNEC_BLOCK + tid;
        if (k < nz) {
            int64_t j = (int64_t)y0 + k;
            s_val[k] = load_value<ENC>(a, j);
            s_col[k] = __builtin_nontemporal_load(a.col + j);
        }
    }
    __syncthreads();

    const int items = nrow + nz;
    const int d0 = tid * IPT < items ? tid * IPT : items;
    const int d1 = d0 + IPT < items ? d0 + IPT : items;
    int rx = (int)merge_search(s_rowend, nrow, y0, nz, d0);
    int ny = d0 - rx;
    double acc[K], head_val[K];
#pragma unroll
    for (int c = 0; c < K; c++) acc[c] = head_val[c] = 0.0;
    int flag = 0, head_row = 0;
#pragma unroll
    for (int u = 0; u < IPT; u++) {
        if (d0 + u < d1) {
            if (rx < nrow && s_rowend[rx] <= y0 + ny) {
                if (!flag) {
                    head_row = x0 + rx;
#pragma unroll
                    for (int c = 0; c < K; c++) head_val[c] = acc[c];
                    flag = 1;
                } else {
                    double *yr = a.y + (int64_t)(x0 + rx) * K;
#pragma unroll
                    for (int c = 0; c < K; c++) yr[c] = acc[c];
                }
#pragma unroll
                for (int c = 0; c < K; c++) acc[c] = 0.0;
                rx++;
            } else {
                const double v = s_val[ny];
                const double *xr = a.x + (int64_t)s_col[ny] * K;
#pragma unroll
                for (int c = 0; c < K; c++) acc[c] += v * xr[c];
                ny++;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < K; c++) {
        double incl;
        double cin = block_seg_carry(flag, acc[c], s_f, s_v, &incl);
        if (flag) a.y[(int64_t)head_row * K + c] = head_val[c] + cin;
        if (tid == VNEC_BLOCK - 1) a.carry[t * K + c] = incl;
    }
}

// VW consecutive right-hand-side entries of one x or y row, held by a lane
// (VW = 2: one 16-byte load or store).
template <int VW> struct rhs_vec { double e[VW]; };
template <int VW> __device__ __forceinline__ rhs_vec<VW> load_rhs(const double *p);
template <> __device__ __forceinline__ rhs_vec<2> load_rhs<2>(const double *p)
{
    const double2 d = *(const double2 *)p;
    return {{d.x, d.y}};
}
__device__ __forceinline__ void store_rhs(double *p, const double (&v)[2])
{
    *(double2 *)p = make_double2(v[0], v[1]);
}

// Segmented scan over merge-path streams of Q lanes each (lane q of a
// stream holds columns [q VW, q VW + VW) of the stream's (flag, tail)):
// Hillis-Steele steps of Q, 2Q, ... lanes within the wavefront, then the
// wavefronts chained in order through LDS.  cin = carry into the stream's
// head row, incl = the stream's inclusive value (the last stream's closes
// the tile).  s_v: VNEC_NWAVES * Q * VW doubles, s_f: VNEC_NWAVES ints.
template <int Q, int VW>
__device__ __forceinline__ void stream_seg_carry(int flag, const double (&tail)[VW], int *s_f, double *s_v,
                                                 double (&cin)[VW], double (&incl)[VW])
{
    constexpr int K = Q * VW;
    const int lane = lane_id(), w = wave_id(), q = lane % Q;
    int f = flag;
    double v[VW];
#pragma unroll
    for (int e = 0; e < VW; e++) v[e] = tail[e];
#pragma unroll
    for (int off = Q; off < VNEC_WAVE; off <<= 1) {
        const int fu = __shfl_up(f, off, VNEC_WAVE);
        double vu[VW];
#pragma unroll
        for (int e = 0; e < VW; e++) vu[e] = __shfl_up(v[e], off, VNEC_WAVE);
        if (lane >= off) {
            if (!f) {
#pragma unroll
                for (int e = 0; e < VW; e++) v[e] = vu[e] + v[e];
            }
            f |= fu;
        }
    }
    if (lane >= VNEC_WAVE - Q) {
#pragma unroll
        for (int e = 0; e < VW; e++) s_v[w * K + q * VW + e] = v[e];
        if (q == 0) s_f[w] = f;
    }
    __syncthreads();
    if (threadIdx.x < Q) {
        double c[VW];
#pragma unroll
        for (int e = 0; e < VW; e++) c[e] = 0.0;
        for (int k = 0; k < VNEC_NWAVES; k++) {
            const int fw = s_f[k];
#pragma unroll
            for (int e = 0; e < VW; e++) {
                const double tw = s_v[k * K + q * VW + e];
                s_v[k * K + q * VW + e] = c[e];  // carry into wavefront k
                c[e] = fw ? tw : c[e] + tw;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < VW; e++) {
        const double cw = s_v[w * K + q * VW + e];
        incl[e] = f ? v[e] : v[e] + cw;
        const double prev = __shfl_up(incl[e], Q, VNEC_WAVE);
        cin[e] = lane < Q ? cw : prev;
    }
}

// Multi-RHS (variant L), grouped lanes: x is [ncols][K], y is [nrows][K].
// Values, columns and row ends are staged in LDS as in k_spmv_mergeK; the
// tile's merge items are then cut into VNEC_BLOCK / Q streams of Q IPT
// items, each walked by a group of Q = K / VW lanes, lane q summing columns
// [q VW, q VW + VW).  A nonzero's K contiguous x entries are thus read by
// one load instruction of the group (8 K bytes in one or two cache lines)
// where k_spmv_mergeK has each lane load its own row, every instruction
// touching 64 lines; rows are written the same way.  The walk decides U
// items (LDS reads only) before issuing their U gathers.  x, y and the
// carries must be 16-byte aligned.
template <int IPT, int K, int ENC, int VW>
__global__ __launch_bounds__(VNEC_BLOCK) void k_spmv_mergeKg(vnec_merge_args a, int64_t grid)
{
    constexpr int T = VNEC_BLOCK * IPT;
    constexpr int Q = K / VW;   // lanes per stream
    constexpr int W = Q * IPT;  // merge items per stream
    constexpr int U = 4;        // items decided before their gathers are issued
    __shared__ double s_val[T];
    __shared__ int s_col[T];
    __shared__ int s_rowend[T];
    __shared__ int s_f[VNEC_NWAVES];
    __shared__ double s_v[VNEC_NWAVES * K];

    const int64_t t = xcd_tile(blockIdx.x, grid);
    if (t >= a.ntiles) return;
    const int2 *co = (const int2 *)a.coords;
    const int2 c0 = co[t], c1 = co[t + 1];
    const int x0 = c0.x, y0 = c0.y;
    const int nrow = c1.x - c0.x, nz = c1.y - c0.y;
    const int tid = threadIdx.x;

    for (int i = tid; i < nrow; i += VNEC_BLOCK) s_rowend[i] = a.ptr[x0 + 1 + i];
#pragma unroll
    for (int u = 0; u < IPT; u++) {
        int k = u * VNEC_BLOCK + tid;
        if (k < nz) {
            int64_t j = (int64_t)y0 + k;
            s_val[k] = load_value<ENC>(a, j);
            s_col[k] = __builtin_nontemporal_load(a.col + j);
        }
    }
    __syncthreads();

    const int q = tid % Q, sid = tid / Q;
    const int items = nrow + nz;
    const int d0 = sid * W < items ? sid * W : items;
    const int d1 = d0 + W < items ? d0 + W : items;
    int rx = (int)merge_search(s_rowend, nrow, y0, nz, d0);
    int ny = d0 - rx;
    int nend = rx < nrow ? s_rowend[rx] : INT_MAX;  // next row end, as a nonzero index
    const double *__restrict__ xq = a.x + q * VW;
    double acc[VW], head[VW];
#pragma unroll
    for (int e = 0; e < VW; e++) acc[e] = head[e] = 0.0;
    int flag = 0, head_row = 0;
    for (int d = d0; d < d1; d += U) {
        int cc[U], er[U];  // er: -2 past the stream, -1 a nonzero, else the row it ends
        double vv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            er[u] = -2;
            cc[u] = 0;
            vv[u] = 0.0;
            if (d + u < d1) {
                if (nend <= y0 + ny) {
                    er[u] = rx++;
                    nend = rx < nrow ? s_rowend[rx] : INT_MAX;
                } else {
                    er[u] = -1;
                    cc[u] = s_col[ny];
                    vv[u] = s_val[ny];
                    ny++;
                }
            }
        }
        rhs_vec<VW> xv[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (er[u] == -1) xv[u] = load_rhs<VW>(xq + (int64_t)cc[u] * K);
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (er[u] == -1) {
#pragma unroll
                for (int e = 0; e < VW; e++) acc[e] += vv[u] * xv[u].e[e];
            } else if (er[u] >= 0) {
                if (!flag) {
                    head_row = x0 + er[u];
#pragma unroll
                    for (int e = 0; e < VW; e++) head[e] = acc[e];
                    flag = 1;
                } else {
                    store_rhs(a.y + (int64_t)(x0 + er[u]) * K + q * VW, acc);
                }
#pragma unroll
                for (int e = 0; e < VW; e++) acc[e] = 0.0;
            }
        }
    }
    double cin[VW], incl[VW];
    stream_seg_carry<Q, VW>(flag, acc, s_f, s_v, cin, incl);
    if (flag) {
#pragma unroll
        for (int e = 0; e < VW; e++) head[e] += cin[e];
        store_rhs(a.y + (int64_t)head_row * K + q * VW, head);
    }
    if (sid == VNEC_BLOCK / Q - 1) store_rhs(a.carry + t * K + q * VW, incl);
}

// Deterministic fix-up: the carry of tile t belongs to row coords[t+1].x.
// One thread per tile boundary and right-hand side: the K threads of a
// boundary are adjacent lanes, so its K carries and y entries are read and
// written as one contiguous piece.  The threads of the first tile of a run
// of tiles carrying into one row sum the run in tile order and add it once
// (the reference's parallel `+=` races when two partitions carry into one
// row, spmv_vnec_L_for_mxv.c:160-164).  Runs longer than FIX_SHORT tiles
// (rows of hundreds of thousands of nonzeros) are summed by the whole
// wavefront instead, 64 carries at a time joined by a fixed butterfly.
// Trailing tiles can be empty (ipt = ceil(M / P) may cover M early); their
// carries belong to row nrows, which does not exist, and are zero.
// KT > 0 fixes K at compile time; KT = 0 takes it from Kr.
#define FIX_SHORT 8
template <typename F, int KT>
__global__ __launch_bounds__(VNEC_BLOCK) void k_carry_fixup(const int2 *__restrict__ co,
                                                            const F *__restrict__ carry, int64_t ntiles,
                                                            F *__restrict__ y, int64_t nrows, int Kr)
{
    const int K = KT > 0 ? KT : Kr;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t t = g / K;        // boundary between tiles t and t + 1
    const int c = (int)(g - t * K);  // right-hand side
    const int64_t nb = ntiles - 1;
    int row = -1;
    bool head = false, longrun = false;
    if (t < nb) {
        row = co[t + 1].x;
        head = row < nrows && (t == 0 || co[t].x != row);
        // runs are contiguous: longer than FIX_SHORT iff tile t + FIX_SHORT still carries into row
        longrun = head && t + FIX_SHORT < nb && co[t + FIX_SHORT + 1].x == row;
    }
    if (head && !longrun) {
        // the run is tiles t .. t + n - 1, n <= FIX_SHORT: the end rows and
        // carries of all FIX_SHORT candidates are loaded together, then the
        // run's carries are summed in tile order
        int er[FIX_SHORT];
        F cv[FIX_SHORT];
#pragma unroll
        for (int i = 0; i < FIX_SHORT; i++) {
            const bool ok = t + i < nb;
            er[i] = i == 0 ? row : (ok ? co[t + i + 1].x : -1);
            cv[i] = ok ? carry[(t + i) * K + c] : (F)0;
        }
        F s = 0;
        bool in = true;
#pragma unroll
        for (int i = 0; i < FIX_SHORT; i++) {
            in = in && er[i] == row;
            if (in) s += cv[i];
        }
        y[(int64_t)row * K + c] += s;
    }
    const int lane = lane_id();
    for (uint64_t m = __ballot(longrun && c == 0); m; m &= m - 1) {
        const int src = __builtin_ctzll(m);
        const int r = __shfl(row, src, VNEC_WAVE);
        const int64_t tr = (g - lane + src) / K;  // lane src's boundary
        for (int cc = 0; cc < K; cc++) {
            F s = 0;
            for (int64_t base = tr;; base += VNEC_WAVE) {
                const int64_t u = base + lane;
                const bool in = u < nb && co[u + 1].x == r;
                if (in) s += carry[u * K + cc];
                if (__ballot(!in)) break;  // the run ends in this batch
            }
            for (int o = VNEC_WAVE / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, VNEC_WAVE);
            if (lane == 0) y[(int64_t)r * K + cc] += s;
        }
    }
}

// ----------------------------------------------------------- dispatch
typedef void (*merge_kernel_t)(vnec_merge_args, int64_t);

#define IPT_LIST(X, ...) X(3, __VA_ARGS__) X(5, __VA_ARGS__) X(7, __VA_ARGS__) X(9, __VA_ARGS__) \
    X(11, __VA_ARGS__) X(13, __VA_ARGS__) X(15, __VA_ARGS__)

static merge_kernel_t pick1(int ipt, bool win, int enc, bool sort)
{
#define CASE1(I, W, E) \
    if (ipt == I && win == W && enc == E && !sort) return k_spmv_merge1<I, W, E>; \
    if (ipt == I && !W && E == 0 && !win && enc == 0 && sort) return k_spmv_merge1<I, false, 0, true>; \
    if (ipt == I && !W && E == VNEC_ENC_F32_K && !win && enc == E && sort) \
        return k_spmv_merge1<I, false, VNEC_ENC_F32_K, true>;
#define ENC3(I, W) CASE1(I, W, 0) CASE1(I, W, 1) CASE1(I, W, 2)
#define WIN2(I, _) ENC3(I, false) ENC3(I, true) CASE1(I, false, VNEC_ENC_F32_K)
    IPT_LIST(WIN2, 0)
    return nullptr;
#undef WIN2
#undef ENC3
#undef CASE1
}

// k >= 4 with 16-byte aligned x and y: grouped lanes, two right-hand sides
// per lane; otherwise (k = 2, where one lane per nonzero is faster, or
// caller pointers that are only 8-byte aligned) the one-lane kernel.
static merge_kernel_t pickK(int ipt, int k, int enc, bool aligned)
{
    const bool grouped = k >= 4 && aligned;
#define CASEK(I, K_, E) \
    if (ipt == I && k == K_ && enc == E) \
        return grouped ? k_spmv_mergeKg<I, K_, E, 2> : k_spmv_mergeK<I, K_, E>;
#define ENCK(I, K_) CASEK(I, K_, 0) CASEK(I, K_, 1) CASEK(I, K_, 2)
#define KS(I, _) ENCK(I, 2) ENCK(I, 4) ENCK(I, 8) ENCK(I, 16)
    if (ipt == 3) { KS(3, 0) }
    if (ipt == 5) { KS(5, 0) }
    if (ipt == 7) { KS(7, 0) }
    return nullptr;
#undef KS
#undef ENCK
#undef CASEK
}

static merge_kernel_t pick(int variant, int ipt, int enc, int nrhs, bool aligned)
{
    if (variant != 0 && variant != 1 && variant != 6) return nullptr;
    if (nrhs == 1) return pick1(ipt, variant == 1, enc, variant == 6);
    if (variant != 0) return nullptr;
    return pickK(ipt, nrhs, enc, aligned);
}

extern "C" int vnec_hip_merge_supported(int variant, int ipt, int encoding, int nrhs)
{
    return pick(variant, ipt, encoding, nrhs, true) != nullptr;
}

extern "C" int vnec_hip_spmv_merge(const vnec_merge_args *a, void *stream)
{
    const bool aligned = (((uintptr_t)a->x | (uintptr_t)a->y | (uintptr_t)a->carry) & 15) == 0;
    merge_kernel_t k = pick(a->variant, a->ipt, a->encoding, a->nrhs, aligned);
    if (!k) return (int)hipErrorInvalidValue;
    if (a->ntiles <= 0) return 0;
    int64_t grid = xcd_grid(a->ntiles);
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(VNEC_BLOCK), 0, (hipStream_t)stream, *a, grid);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_carry_fixup(const int32_t *coords, const double *carry, int64_t ntiles,
                                    double *y, int64_t nrows, int nrhs, void *stream)
{
    if (ntiles <= 1) return 0;
    /* one thread per boundary and right-hand side */
    int64_t grid = ((ntiles - 1) * (int64_t)(nrhs > 0 ? nrhs : 1) + VNEC_BLOCK - 1) / VNEC_BLOCK;
    void (*k)(const int2 *, const double *, int64_t, double *, int64_t, int) =
        nrhs == 1 ? k_carry_fixup<double, 1> : nrhs == 2 ? k_carry_fixup<double, 2>
        : nrhs == 4 ? k_carry_fixup<double, 4> : nrhs == 8 ? k_carry_fixup<double, 8>
        : nrhs == 16 ? k_carry_fixup<double, 16> : k_carry_fixup<double, 0>;
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(VNEC_BLOCK), 0, (hipStream_t)stream,
                       (const int2 *)coords, carry, ntiles, y, nrows, nrhs);
    return (int)hipGetLastError();
}

extern "C" int vnec_hip_carry_fixup_f32(const int32_t *coords, const float *carry, int64_t ntiles,
                                        float *y, int64_t nrows, void *stream)
{
    if (ntiles <= 1) return 0;
    int64_t grid = (ntiles - 1 + VNEC_BLOCK - 1) / VNEC_BLOCK;
    hipLaunchKernelGGL((k_carry_fixup<float, 1>), dim3((unsigned)grid), dim3(VNEC_BLOCK), 0, (hipStream_t)stream,
                       (const int2 *)coords, carry, ntiles, y, nrows, 1);
    return (int)hipGetLastError();
}
