"""Captured-graph SpMV (vnec_spmv_graph_ext) and SpB_mxv on device-resident
vectors: every replay gives the bits of the direct launch
(vnec_spmv_dev_ext), with eager work and synchronisations between the
replays, for every plan shape; x's contents may change between replays;
a plan keeps graphs for four pointer pairs and recaptures a fifth."""
import numpy as np
import pytest

import oracle as O
from vnec_amd import matgen as g
from vnec_amd import spb

pytestmark = pytest.mark.gpu

TOL = 1e-12

SHAPES = [("L", 7, 0, 1, "fp64", 64), ("L", 11, -1, 1, "fp64", 64), ("C", 9, 0, 1, "fp64", 64),
          ("C", 7, -2, 1, "fp64", 64), ("E", 5, 0, 1, "fp64", 64), ("S", 3, -1, 1, "fp64", 64),
          ("S", 3, -1, 1, "hilo", 64), ("D", 7, -1, 1, "fp64", 8), ("H", 5, 0, 1, "fp64", 4096),
          ("L", 5, -1, 8, "fp64", 64), ("L", 7, 0, 1, "hi", 64),
          ("C", 9, -2, 1, "fp64", 1), ("C", 7, 0, 1, "fp64", 1024),  # relabelled; wide tiles
          ("C", 7, -1, 1, "fp64", 3072)]  # gathers issued together


def _sync():
    spb.L.check(spb.L.load().vnec_sync_ext(), "sync")


@pytest.mark.parametrize("variant,ipt,grid,k,enc,lanes", SHAPES,
                         ids=[f"{v}{i}_g{gr}_k{k}_{e}_l{ln}" for v, i, gr, k, e, ln in SHAPES])
def test_graph_replay_is_bit_identical_to_direct_launch(variant, ipt, grid, k, enc, lanes):
    m = g.rmat(14, 16, seed=91)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    p = spb.Plan(A, variant, ipt, lanes=lanes, encoding=enc, nrhs=k, grid=grid)
    rng = np.random.default_rng(7)
    X = rng.uniform(-1, 1, m.ncols * k)
    dx = spb.DeviceArray.from_array(X)
    dy_direct = spb.DeviceArray(8 * m.nrows * k)
    dy_graph = spb.DeviceArray.from_array(np.full(m.nrows * k, np.nan))
    p.spmv_dev(dx.ptr, dy_direct.ptr)
    _sync()
    want = dy_direct.to_array(np.float64, m.nrows * k)
    for rep in range(3):
        p.spmv_graph(dx.ptr, dy_graph.ptr)
        _sync()
        assert np.array_equal(dy_graph.to_array(np.float64, m.nrows * k), want), rep
        p.spmv_dev(dx.ptr, dy_direct.ptr)  # eager work between replays
        _sync()
    # new contents of x at the captured address: the replay reads them
    X2 = rng.uniform(-1, 1, m.ncols * k)
    spb.L.check(spb.L.load().vnec_memcpy_h2d_ext(dx.ptr, X2.ctypes.data, X2.nbytes), "h2d")
    p.spmv_dev(dx.ptr, dy_direct.ptr)
    _sync()
    want2 = dy_direct.to_array(np.float64, m.nrows * k)
    p.spmv_graph(dx.ptr, dy_graph.ptr)
    _sync()
    assert np.array_equal(dy_graph.to_array(np.float64, m.nrows * k), want2)
    if enc == "fp64":
        ax, _ = O.spmv_abs(m.ptr, m.col, m.val, X2[::k])
        assert O.check_rows(want2[::k], O.spmv_serial(m.ptr, m.col, m.val, X2[::k]), ax, TOL)[0]
    p.free()
    A.free()


def test_graph_pattern_slices_and_pointer_pairs():
    """Variant G (with an irregular-row sub-plan) and more pointer pairs than
    the plan keeps graphs for: each pair replays its own capture."""
    m = g.stencil27(24, seed=92)
    # one irregular row (a column outside the 27-point pattern) -> sub-plan
    ptr, col, val = m.ptr.copy(), m.col.copy(), m.val.copy()
    r = 5000
    col[ptr[r]] = (col[ptr[r]] + 777) % m.ncols
    order = np.argsort(col[ptr[r]:ptr[r + 1]], kind="stable")
    col[ptr[r]:ptr[r + 1]] = col[ptr[r]:ptr[r + 1]][order]
    val[ptr[r]:ptr[r + 1]] = val[ptr[r]:ptr[r + 1]][order]
    A = spb.Matrix.from_csr(m.nrows, m.ncols, ptr, col, val)
    p = spb.Plan(A, "G", 8)
    xs = [spb.DeviceArray.from_array(np.random.default_rng(s).uniform(-1, 1, m.ncols))
          for s in range(6)]
    ys = [spb.DeviceArray(8 * m.nrows) for _ in range(6)]
    yd = spb.DeviceArray(8 * m.nrows)
    for rnd in range(2):
        for i in range(6):  # six pairs through four graph slots
            p.spmv_graph(xs[i].ptr, ys[i].ptr)
            p.spmv_dev(xs[i].ptr, yd.ptr)
            _sync()
            assert np.array_equal(ys[i].to_array(np.float64, m.nrows),
                                  yd.to_array(np.float64, m.nrows)), (rnd, i)
    x0 = xs[0].to_array(np.float64, m.ncols)
    ax, _ = O.spmv_abs(ptr, col, val, x0)
    assert O.check_rows(ys[0].to_array(np.float64, m.nrows), O.spmv_serial(ptr, col, val, x0), ax,
                        TOL)[0]
    p.free()
    A.free()


def test_mxv_on_resident_vectors_replays_the_same_bits():
    """SpB_mxv with device-resident x and y (the reference harness's repeated
    SpMV, test_spmv_vnec.cpp:205-214): the same bits call after call, the
    bits of host vectors, and new x values are picked up."""
    m = g.rmat(13, 16, seed=93)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    x = spb.Vector.from_array(m.x)
    y = spb.Vector(m.nrows)
    spb.mxv(y, A, x)  # host vectors: direct launch
    first = y.to_array()
    x.resident(True)
    y.resident(True)
    for _ in range(4):
        spb.mxv(y, A, x)
        assert np.array_equal(y.to_array(), first)
    x.set(3, 0.5)  # host write: the device copy is refreshed before the replay
    xm = m.x.copy()
    xm[3] = 0.5
    spb.mxv(y, A, x)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, xm)
    assert O.check_rows(y.to_array(), O.spmv_serial(m.ptr, m.col, m.val, xm), ax, TOL)[0]
    x.resident(False)
    y.resident(False)
    spb.mxv(y, A, x)
    got = y.to_array()
    x.resident(True)
    y.resident(True)
    spb.mxv(y, A, x)
    assert np.array_equal(y.to_array(), got)
