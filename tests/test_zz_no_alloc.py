"""vnec_csr_spmv_noalloc_ext: the merge-path CSR SpMV that allocates no
device memory (tile coordinates and carries in a fixed device array) gives
the bits of vnec_spmv_dev_ext on an L plan at 7 items per thread with one
workgroup per tile, call after call, on inputs of decreasing size."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from vnec_amd import _lib as L
from vnec_amd import matgen as g
from vnec_amd import spb

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _noalloc(lib, m, ptr, col, val, x, y):
    rc = lib.vnec_csr_spmv_noalloc_ext(m.nrows, m.ncols, m.nnz, ptr.ptr, col.ptr, val.ptr, x.ptr,
                                       y.ptr, None)
    assert rc == 0, L.last_error()
    L.check(lib.vnec_sync_ext(), "sync")
    return y.to_array(np.float64, m.nrows)


def _plan(m, dx):
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    p = spb.Plan(A, "L", 7, grid=-1)
    info = p.info()
    assert (info.variant, info.ipt, info.grid) == ("L", 7, -1)
    dy = spb.DeviceArray(8 * m.nrows)
    p.spmv_dev(dx.ptr, dy.ptr)
    L.check(L.load().vnec_sync_ext(), "sync")
    y = dy.to_array(np.float64, m.nrows)
    p.free()
    A.free()
    return y


def test_noalloc_spmv_matches_the_plan_call_after_call():
    lib = spb.init()
    # the first matrix has rows longer than a tile (carries across tiles); the
    # second is smaller, so stale coordinates and carries of the first call
    # lie beyond its tiles in the shared array
    for m in (g.rmat(15, 16, 0.57, 0.19, 0.19, seed=71), g.stencil27(12, seed=72)):
        print(f"no-alloc SpMV: {m.name} nrows {m.nrows} nnz {m.nnz}", flush=True)
        assert int(np.diff(m.ptr).max()) > 0
        ptr = spb.DeviceArray.from_array(m.ptr.astype(np.int32))
        col = spb.DeviceArray.from_array(m.col)
        val = spb.DeviceArray.from_array(m.val)
        dx = spb.DeviceArray.from_array(m.x)
        dy = spb.DeviceArray.from_array(np.full(m.nrows, np.nan))  # every row is written
        y = _noalloc(lib, m, ptr, col, val, dx, dy)
        y_plan = _plan(m, dx)
        assert np.array_equal(y, y_plan)
        y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
        ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
        ok, worst = O.check_rows(y, y0, ax, TOL)
        assert ok, worst
        # the same call again leaves the same bits
        assert np.array_equal(_noalloc(lib, m, ptr, col, val, dx, dy), y_plan)


def test_noalloc_spmv_rejects_inputs_beyond_its_scratch():
    lib = spb.init()
    m = g.stencil27(6, seed=73)
    ptr = spb.DeviceArray.from_array(m.ptr.astype(np.int32))
    col = spb.DeviceArray.from_array(m.col)
    val = spb.DeviceArray.from_array(m.val)
    dx = spb.DeviceArray.from_array(m.x)
    dy = spb.DeviceArray.from_array(np.zeros(m.nrows))
    # 65536 tiles of 1792 items hold nrows + nnz <= 117,440,512
    rc = lib.vnec_csr_spmv_noalloc_ext(60_000_000, 60_000_000, 57_440_513, ptr.ptr, col.ptr, val.ptr,
                                       dx.ptr, dy.ptr, None)
    assert rc == 1 and "scratch holds 65536" in L.last_error()
    # nothing ran: y is untouched, and a valid call still works
    L.check(lib.vnec_sync_ext(), "sync")
    assert np.array_equal(dy.to_array(np.float64, m.nrows), np.zeros(m.nrows))
    y = _noalloc(lib, m, ptr, col, val, dx, dy)
    assert np.array_equal(y, _plan(m, dx))
