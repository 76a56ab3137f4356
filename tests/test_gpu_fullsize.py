"""The benchmark configurations at their full sizes, through the C-ABI, each
checked row by row against the oracle: |y - y_ref| <= 1e-12 (|A||x|)_i with
y_ref the long-double row sums (2^-23 for FP32 hi-only slots), and every
plan run twice with identical bits.

* R-MAT scale 18, edge factor 16, seed 1: the learned plan and the plain
  merge-path kernel;
* 27-point stencil 128^3: variant G with FP64, FP32 hi + lo and FP32 hi
  slots (the hi-only error must exceed 1e-12: the low bits are really gone);
* skew (R-MAT 22 + 8 dense rows of 500k columns): the learned plan, whose
  dense rows span hundreds of tiles;
* the 512-member batch, k = 1 and k = 8 (the grouped multi-RHS kernel),
  against the block-diagonal CSR with member columns made global;
* R-MAT 18 and the stencil once more, against the y of the reference
  library compiled from its own sources (its IRD dispatch, VNEC-D or -S).
"""
import numpy as np
import pytest

import oracle as O
from vnec_amd import matgen as g
from vnec_amd import spb

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if spb.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    spb.init()


def _run_twice(p, x, n_out):
    dx = spb.DeviceArray.from_array(x)
    dy = spb.DeviceArray(8 * n_out)
    p.spmv_dev(dx.ptr, dy.ptr)
    y = dy.to_array(np.float64, n_out)
    p.spmv_dev(dx.ptr, dy.ptr)
    y2 = dy.to_array(np.float64, n_out)
    dx.free()
    dy.free()
    assert np.array_equal(y, y2), "repeated SpMV changed bits"
    return y


def _gate(m, y, tol=TOL):
    ax, yl = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ok, worst = O.check_rows(y, yl, ax, tol)
    assert ok, worst
    return worst


def test_rmat18_full_size():
    m = g.rmat(18, 16, 0.57, 0.19, 0.19, seed=1)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    for mode, ipt in (("learned", 7), ("L", 7)):
        print(f"rmat18 nnz {m.nnz} plan {mode}", flush=True)
        p = spb.Plan(A, mode, ipt)
        _gate(m, _run_twice(p, m.x, m.nrows))
        p.free()


_STENCIL = []


def _stencil():
    if not _STENCIL:
        _STENCIL.append(g.stencil27(128, seed=1))
    return _STENCIL[0]


@pytest.mark.parametrize("enc,tol", [("fp64", TOL), ("hilo", TOL), ("hi", 2.0 ** -23)])
def test_stencil128_variant_g(enc, tol):
    m = _stencil()
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    print(f"stencil 128^3 variant G encoding {enc}", flush=True)
    p = spb.Plan(A, "G", encoding=enc)
    info = p.info()
    assert info.variant == "G" and info.encoding == enc
    worst = _gate(m, _run_twice(p, m.x, m.nrows), tol)
    if enc == "hi":
        assert worst > 1e-12
    p.free()


def _avx512():
    try:
        return "avx512f" in open("/proc/cpuinfo").read()
    except OSError:
        return False


@pytest.mark.parametrize("name", ["rmat18", "stencil128"])
def test_full_size_against_the_reference_library(name):
    """The learned plan's y at the benchmark sizes against the y of the
    reference library itself (compiled from its sources; its dispatch: IRD,
    then VNEC-D or VNEC-S at P = 36, L = 8), row by row:
    |y - y_ref| <= 1e-12 (|A||x|)_i."""
    if not (O.ref_available() and _avx512()):
        pytest.skip("the reference library needs its build and an AVX-512 host")
    m = g.rmat(18, 16, 0.57, 0.19, 0.19, seed=1) if name == "rmat18" else _stencil()
    R = O.RefMatrix(m.nrows, m.ncols, m.ptr, m.col, m.val)
    kind = "D" if R.ird() == 0 else "S"
    print(f"{name} nnz {m.nnz}: reference VNEC-{kind}", flush=True)
    plan = R.build_d() if kind == "D" else R.build_s()
    y_ref = R.spmv(kind, plan, m.x)
    R.free(kind, plan)
    R.close()
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    p = spb.Plan(A, "learned")
    y = _run_twice(p, m.x, m.nrows)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ok, worst = O.check_rows(y, y_ref, ax, TOL)
    assert ok, (name, kind, worst)
    p.free()


def test_skew_full_size():
    m = g.skew(22, 16, 8, 500_000, seed=1)
    assert int(np.diff(m.ptr).max()) >= 500_000
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    print(f"skew nnz {m.nnz} learned plan", flush=True)
    p = spb.Plan(A, "learned")
    _gate(m, _run_twice(p, m.x, m.nrows))
    p.free()


@pytest.mark.parametrize("k,variant,ipt,lanes", [(1, "L", 11, 64), (1, "C", 9, 1024), (8, "L", 5, 64)])
def test_batch512_full_size(k, variant, ipt, lanes):
    """The 512-matrix batch at full size: L11 and the autotuner's pick for
    it (C9 in 512-thread tiles, column-sorted) with per-tile carries, and
    the grouped multi-RHS kernel at k = 8, against the oracle per member."""
    m = g.batch(512, 1000, 200_000, 4.0, 64.0, nrhs=k, seed=1)
    A = spb.Matrix.from_batch(m.row_off, m.col_off, m.ptr, m.col, m.val)
    print(f"batch512 nnz {m.nnz} k {k} plan {variant}{ipt} lanes {lanes}", flush=True)
    p = spb.Plan(A, variant, ipt, nrhs=k, grid=-1, lanes=lanes)
    if k == 1:  # hundreds of tiles per workgroup: the strided persistent launch
        assert p.info().grid == -2
    y = _run_twice(p, m.x, m.nrows * k)
    p.free()
    col = m.col + np.repeat(m.col_off[:-1], np.diff(m.ptr[m.row_off])).astype(np.int32)
    for c in range(k):
        ax, yl = O.spmv_abs(m.ptr, col, m.val, np.ascontiguousarray(m.x[c::k]))
        ok, worst = O.check_rows(y[c::k], yl, ax, TOL)
        assert ok, (c, worst)
