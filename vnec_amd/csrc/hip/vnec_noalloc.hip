// vnec_noalloc.hip — merge-path CSR SpMV y = A x on caller device arrays
// that allocates nothing: for callers that cannot call hipMalloc (code
// under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// It runs what an FP64 L plan at 7 items per thread with one workgroup per
// tile runs (vnec_plan_build_ext + vnec_spmv_dev_ext): the merge-path
// partition of the diagonals min(t * 1792, M) (M = nrows + nnz), the
// one-workgroup-per-tile merge kernel, and the ordered carry fix-up.  The
// tile coordinates and per-tile carries that the plan keeps in memory it
// allocates live here in `scratch`, a file-scope device array of
// NOALLOC_SCRATCH_BYTES (about 1 MiB) that each GPU the library runs on
// holds until the process ends; that bounds the launch to
// VNEC_NOALLOC_MAX_TILES tiles (nrows + nnz <= 117,440,512).
//
// Every call on one GPU shares that GPU's array and every tile rewrites its
// coordinates and carry before they are read, so nothing needs clearing
// between calls; calls that overlap on one GPU (two host threads, or two
// streams) race on the array.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>

#include "vnec_ext.h"
#include "vnec_hip.h"

extern "C" SpB_Info vnec_fail(const char *fmt, ...);
extern "C" hipStream_t vnec_stream(void);

#define NOALLOC_IPT 7
#define NOALLOC_TILE_ITEMS (256 * NOALLOC_IPT)
// coordinates: one int2 per tile boundary; carries: one double per tile
#define NOALLOC_COORD_BYTES ((((int64_t)VNEC_NOALLOC_MAX_TILES + 1) * 8 + 255) / 256 * 256)
#define NOALLOC_SCRATCH_BYTES (NOALLOC_COORD_BYTES + (int64_t)VNEC_NOALLOC_MAX_TILES * 8)
#define NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Per-device address of `scratch`, resolved once per device with that
// device current.  In host code the array's name is not a device address;
// only these pointers are passed to kernels.
static void *g_scratch_dev[NOALLOC_MAX_DEVICES];

static int scratch_address(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return (int)hipErrorInvalidDevice;
    void *p = __atomic_load_n(&g_scratch_dev[dev], __ATOMIC_ACQUIRE);
    if (!p) {
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        __atomic_store_n(&g_scratch_dev[dev], p, __ATOMIC_RELEASE);
    }
    *out = p;
    return (int)hipSuccess;
}

extern "C" SpB_Info vnec_csr_spmv_noalloc_ext(int64_t nrows, int64_t ncols, int64_t nnz, const int32_t *ptr_dev,
                                              const int32_t *col_dev, const double *val_dev, const double *x_dev,
                                              double *y_dev, void *stream)
{
    if (nrows < 0 || ncols < 0 || nnz < 0)
        return vnec_fail("vnec_csr_spmv_noalloc_ext: negative dimension");
    if (nrows == 0 && nnz > 0) return vnec_fail("vnec_csr_spmv_noalloc_ext: %lld nonzeros in no rows", (long long)nnz);
    const int64_t M = nrows + nnz;
    if (M >= INT32_MAX) return vnec_fail("vnec_csr_spmv_noalloc_ext: nrows + nnz must stay below 2^31");
    const int64_t ntiles = (M + NOALLOC_TILE_ITEMS - 1) / NOALLOC_TILE_ITEMS;
    if (ntiles > VNEC_NOALLOC_MAX_TILES)
        return vnec_fail("vnec_csr_spmv_noalloc_ext: %lld rows + %lld nonzeros need %lld tiles; its scratch "
                         "holds %d (nrows + nnz <= %lld)",
                         (long long)nrows, (long long)nnz, (long long)ntiles, VNEC_NOALLOC_MAX_TILES,
                         (long long)VNEC_NOALLOC_MAX_TILES * NOALLOC_TILE_ITEMS);
    if (ntiles == 0) return SpB_SUCCESS;
    if (!ptr_dev || !y_dev || (nnz > 0 && (!col_dev || !val_dev || !x_dev)))
        return vnec_fail("vnec_csr_spmv_noalloc_ext: NULL array");
    hipStream_t s = stream ? (hipStream_t)stream : vnec_stream();
    void *base = NULL;
    int e = scratch_address(&base);
    if (e != (int)hipSuccess)
        return vnec_fail("vnec_csr_spmv_noalloc_ext: scratch address: %s", hipGetErrorString((hipError_t)e));
    int32_t *coords = (int32_t *)base;
    double *carry = (double *)((unsigned char *)base + NOALLOC_COORD_BYTES);

    e = vnec_hip_partition(ptr_dev, nrows, nnz, ntiles, 0, coords, s);
    if (e) return vnec_fail("vnec_csr_spmv_noalloc_ext: partition: %s", hipGetErrorString((hipError_t)e));
    vnec_merge_args m = {};
    m.ptr = ptr_dev;
    m.col = col_dev;
    m.val = val_dev;
    m.coords = coords;
    m.ntiles = ntiles;
    m.x = x_dev;
    m.y = y_dev;
    m.carry = carry;
    m.nrows = nrows;
    m.ipt = NOALLOC_IPT;
    m.variant = VNEC_VAR_L;
    m.encoding = VNEC_ENC_FP64;
    m.nrhs = 1;
    e = vnec_hip_spmv_merge(&m, s);
    if (e) return vnec_fail("vnec_csr_spmv_noalloc_ext: spmv kernel: %s", hipGetErrorString((hipError_t)e));
    e = vnec_hip_carry_fixup(coords, carry, ntiles, y_dev, nrows, 1, s);
    if (e) return vnec_fail("vnec_csr_spmv_noalloc_ext: carry fix-up: %s", hipGetErrorString((hipError_t)e));
    return SpB_SUCCESS;
}
