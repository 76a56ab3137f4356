"""GPU parity of the FP32 API (Build_VNEC_{D,S}_FP32, SpMV_VNEC_{L,D,S}_FP32,
IRD_VNEC(A, SpB_FP32), SpB_mxv on FP32 objects, plans with encoding FP32)
through the C-ABI.

Integer structures at the reference's FP32 parameters (P = 36, L = 16)
must equal the reference's own FP32 builders bit for bit (fixtures
`<name>_f32.npz`).  FP32 results are gated per row against the harness's
FP32 serial SpMV (the oracle) by the reference's tolerance 1e-4
(test_spmv_vnec.cpp:132) relative to (|A||x|)_i, widened to the FP32
summation bound 2 (len_i + 1) 2^-24 on rows long enough for that bound to
exceed it (oracle.check_rows_f32).
"""
import numpy as np
import pytest

import oracle as O
from conftest import golden_names, load_golden
from vnec_amd import matgen as g
from vnec_amd import spb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if spb.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    spb.init()


def _f32(m):
    return (m.ptr, m.col, np.asarray(m.val, np.float32), np.asarray(m.x, np.float32))


def _gate(y, ptr, col, val, x):
    y0 = O.spmv_serial_f32(ptr, col, val, x)
    ax, _ = O.spmv_abs(ptr, col, np.asarray(val, np.float64), np.asarray(x, np.float64))
    return O.check_rows_f32(y, y0, ptr, ax)


def _cases():
    out = [("rmat12", g.rmat(12, 16, seed=21)),
           ("stencil16", g.stencil27(16, seed=22)),
           ("powerlaw", g.powerlaw(6000, 1.7, 2, 5000, seed=23)),
           ("banded", g.banded(5000, 40, seed=24))]
    n = 3000
    lens = np.zeros(n, np.int64)
    lens[::3] = 5
    lens[1500] = 2500  # one row spanning many tiles, empty rows around it
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.sort(np.random.default_rng(i).choice(n, int(l), replace=False))
                          for i, l in enumerate(lens) if l] + [np.zeros(0, np.int64)]).astype(np.int32)
    rng = np.random.default_rng(5)
    out.append(("longrow", g.Csr(n, n, ptr, col, rng.uniform(-1, 1, len(col)),
                                 rng.uniform(-1, 1, n), name="longrow")))
    return out


CASES = _cases()


@pytest.mark.parametrize("name,m", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("variant,ipt,lanes", [("L", 3, 64), ("L", 7, 64), ("L", 15, 64),
                                                ("C", 3, 64), ("C", 7, 64), ("C", 11, 64),
                                                ("C", 5, 1024), ("C", 9, 1024), ("C", 5, 2048),
                                                ("C", 7, 3072),
                                                ("S", 3, 64), ("D", 5, 16), ("D", 7, 8),
                                                ("D", 9, 4)])
def test_fp32_plans_match_the_fp32_serial_oracle(name, m, variant, ipt, lanes):
    ptr, col, val, x = _f32(m)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, ptr, col, val, dtype=np.float32)
    p = spb.Plan(A, variant, ipt, lanes=lanes, grid=-1)
    assert p.info().encoding == "fp32"
    xv = spb.Vector.from_array(x, np.float32)
    yv = spb.Vector(m.nrows, np.float32)
    p.spmv(yv, xv)
    y = yv.to_array()
    assert y.dtype == np.float32
    ok, worst = _gate(y, ptr, col, val, x)
    assert ok, (name, variant, worst)
    # a second SpMV over the same plan gives the same bits (deterministic fix-up)
    p.spmv(yv, xv)
    assert np.array_equal(yv.to_array().view(np.uint32), y.view(np.uint32))
    p.free()


def test_fp32_empty_and_tiny():
    A = spb.Matrix.from_csr(5, 4, np.zeros(6, np.int64), np.zeros(0, np.int32),
                            np.zeros(0, np.float32), dtype=np.float32)
    y = spb.Vector.from_array(np.full(5, 7.0, np.float32), np.float32)
    spb.mxv(y, A, spb.Vector.from_array(np.ones(4, np.float32), np.float32))
    assert np.array_equal(y.to_array(), np.zeros(5, np.float32))
    B = spb.Matrix.from_csr(1, 1, [0, 1], [0], [3.0], dtype=np.float32)
    for kind in ("L", "D", "S"):
        plan = spb.build_vnec(kind, B)
        yv = spb.Vector(1, np.float32)
        assert spb.spmv_vnec(kind, yv, B, spb.Vector.from_array([2.0], np.float32), plan) == ""
        assert yv.to_array()[0] == np.float32(6.0)
        spb.free_vnec(kind, plan, fp32=True)


@pytest.mark.parametrize("name", golden_names(fp32=True))
def test_fp32_reference_api_bit_exact_structures_and_gated_results(name):
    d = load_golden(name)
    n = int(d["nrows"])
    ptr, col, val, x = d["ptr"], d["col"], d["val"], d["x"]
    A = spb.Matrix.from_csr(n, int(d["ncols"]), ptr, col, val, dtype=np.float32)
    xv = spb.Vector.from_array(x, np.float32)
    assert spb.ird_vnec(A) == int(d["ird"])
    for kind in ("L", "D", "S"):
        plan = spb.build_vnec(kind, A)
        y = spb.Vector(n, np.float32)
        assert spb.spmv_vnec(kind, y, A, xv, plan) == ""
        yy = y.to_array()
        ok, worst = _gate(yy, ptr, col, val, x)
        assert ok, (kind, worst)
        if "y_" + kind in d:  # the reference's own FP32 kernel, same gate
            ax, _ = O.spmv_abs(ptr, col, val.astype(np.float64), x.astype(np.float64))
            assert O.check_rows_f32(yy, d["y_" + kind], ptr, ax)[0], kind
        p = plan.contents
        nz = int(ptr[-1])
        if kind == "D":
            co = np.array([[p.thread_coord_start[t].x, p.thread_coord_start[t].y,
                            p.thread_coord_end[t].x, p.thread_coord_end[t].y] for t in range(36)])
            assert np.array_equal(co, d["coords"])
            if nz:
                assert np.array_equal(np.ctypeslib.as_array(p.ecr_indices, (nz,)), d["ecr"])
            assert np.array_equal(np.ctypeslib.as_array(p.v_row_ptr, (n + 1,)), d["v_row_ptr"])
            G = len(d["col_start"])
            if G:
                assert np.array_equal(np.ctypeslib.as_array(p.col_start, (G,)), d["col_start"])
                # val_align: the reference's G * 16 filled entries
                assert np.array_equal(np.ctypeslib.as_array(p.val_align, (G * 16,)),
                                      d["val_align"][:G * 16])
            assert np.array_equal(np.ctypeslib.as_array(p.NEC_NUM, (36,)), d["nec"])
        if kind == "S" and "s_tasks" in d:
            assert np.array_equal(np.ctypeslib.as_array(p.NEC_NUM, (36,)), d["s_nec"])
            assert np.array_equal(np.ctypeslib.as_array(p.spvv16_len, (36,)), d["s_spvv_len"])
            rm = p.reorder_mat
            ts = np.ctypeslib.as_array(rm.task_start, (36,))
            te = np.ctypeslib.as_array(rm.task_end, (36,))
            assert np.array_equal(np.c_[ts, te], d["s_task_se"])
            tasks = np.concatenate([np.ctypeslib.as_array(p.tasks[t], (int(te[t] - ts[t]),))
                                    for t in range(36) if te[t] > ts[t]] or [np.zeros(0, np.int32)])
            assert np.array_equal(tasks, d["s_tasks"])
            assert np.array_equal(np.ctypeslib.as_array(rm.row_begin, (n,)), d["s_row_begin"])
            assert np.array_equal(np.ctypeslib.as_array(rm.row_end, (n,)), d["s_row_end"])
            if nz:
                assert np.array_equal(np.ctypeslib.as_array(rm.col, (nz,)), d["s_s_col"])
                assert np.array_equal(np.ctypeslib.as_array(rm.nnz, (nz,)), d["s_s_val"])
        spb.free_vnec(kind, plan, fp32=True)
    y = spb.Vector(n, np.float32)
    spb.mxv(y, A, xv)
    assert _gate(y.to_array(), ptr, col, val, x)[0]


def test_fp32_d_layout_and_ird_at_l16_match_the_oracle():
    m = g.banded(4000, 9, seed=31)
    ptr, col, val, _ = _f32(m)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, ptr, col, val, dtype=np.float32)
    vrp, cs, va = spb.d_layout(A, 36, 16)
    ref = O.d_build(ptr, col, val.astype(np.float64), 36, 16)
    assert np.array_equal(vrp, ref.v_row_ptr) and np.array_equal(cs, ref.col_start)
    assert np.array_equal(va, ref.val_align)
    assert spb.ird_vnec(A) == O.ird(ptr, col, 36, 16, 0.437)[0]
    # the FP64 decision of the same structure uses L = 8, 0.477
    assert spb.ird_vnec(A, fp32=False) == O.ird(ptr, col, 36, 8, 0.477)[0]


def test_fp32_entry_points_refuse_mixed_types():
    m = g.stencil27(6, seed=3)
    A32 = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val, dtype=np.float32)
    A64 = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    x32, y32 = spb.Vector(m.ncols, np.float32), spb.Vector(m.nrows, np.float32)
    x64, y64 = spb.Vector(m.ncols), spb.Vector(m.nrows)
    plan = spb.build_vnec("L", A32)
    assert "FP64" in spb.spmv_vnec("L", y64, A32, x64, plan, fp32=False)
    assert "FP32 vectors" in spb.spmv_vnec("L", y64, A32, x64, plan, fp32=True)
    assert spb.spmv_vnec("L", y32, A32, x32, plan) == ""
    spb.free_vnec("L", plan)
    with pytest.raises(RuntimeError, match="VNEC_ENC_FP32"):
        spb.Plan(A32, "L", 7, encoding="fp64")
    with pytest.raises(RuntimeError, match="SpB_FP32"):
        spb.Plan(A64, "L", 7, encoding="fp32")
    with pytest.raises(RuntimeError, match="variants L, C, S, D and G"):
        spb.Plan(A32, "E", 7)
    with pytest.raises(RuntimeError, match="relabelled columns"):
        spb.Plan(A32, "C", 7, lanes=1)
    with pytest.raises(RuntimeError, match="wide tiles"):
        spb.Plan(A32, "C", 13, lanes=1024)


@pytest.mark.parametrize("name,m", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("variant,ipt", [("L", 3), ("L", 7), ("L", 11), ("C", 7), ("C", 9)])
def test_fp32_strided_launch_is_bit_identical_to_one_workgroup_per_tile(name, m, variant, ipt):
    """The FP32 API's merge-path plans (CSR-ordered or column-sorted tiles)
    with per-tile carries from the strided persistent launch (grid -2): the
    tiles, carries and fix-up of grid -1, so the same bits, within the FP32
    gate."""
    ptr, col, val, x = _f32(m)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, ptr, col, val, dtype=np.float32)
    out = []
    for grid in (-1, -2):
        print(f"case {name} FP32 {variant}{ipt} grid {grid}", flush=True)
        p = spb.Plan(A, variant, ipt, grid=grid)
        assert p.info().grid == grid and p.info().encoding == "fp32"
        yv = spb.Vector(m.nrows, np.float32)
        p.spmv(yv, spb.Vector.from_array(x, np.float32))
        out.append(yv.to_array())
        p.free()
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))
    ok, worst = _gate(out[1], ptr, col, val, x)
    assert ok, worst


def test_fp32_selection_takes_pattern_slices_where_the_rule_holds():
    """FP32 matrices in a selection mode (what SpB_mxv uses): pattern slices
    for a stencil of more than 2^20 nonzeros, column-sorted tiles where the
    column-reuse rule holds (R-MAT 17), 64-row slices or merge-path tiles
    elsewhere; results within the FP32 gate."""
    for m, want in ((g.stencil27(40, seed=31), "G"), (g.rmat(17, 16, seed=71), "C"),
                    (g.rmat(14, 16, seed=32), None)):
        ptr, col, val, x = _f32(m)
        A = spb.Matrix.from_csr(m.nrows, m.ncols, ptr, col, val, dtype=np.float32)
        p = spb.Plan(A, "learned")
        info = p.info()
        assert info.encoding == "fp32"
        if want:
            assert info.variant == want
        else:
            assert info.variant in ("S", "L")
        yv = spb.Vector(m.nrows, np.float32)
        p.spmv(yv, spb.Vector.from_array(x, np.float32))
        ok, worst = _gate(yv.to_array(), ptr, col, val, x)
        assert ok, (info, worst)
        y2 = spb.Vector(m.nrows, np.float32)
        spb.mxv(y2, A, spb.Vector.from_array(x, np.float32))
        assert np.array_equal(y2.to_array().view(np.uint32), yv.to_array().view(np.uint32))
        p.free()


def test_fp32_reference_l_entry_takes_column_sorted_tiles_on_column_reuse():
    """Build_VNEC_L on an FP32 matrix whose columns are reused (R-MAT 17,
    the column-reuse rule) runs column-sorted tiles with per-tile carries;
    SpMV_VNEC_L_FP32 stays within the FP32 gate and repeats its bits."""
    m = g.rmat(17, 16, seed=71)
    ptr, col, val, x = _f32(m)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, ptr, col, val, dtype=np.float32)
    plan = spb.build_vnec("L", A)
    info = spb.reference_plan_info(plan)
    assert info.variant == "C" and info.encoding == "fp32" and info.grid in (-1, -2)
    xv = spb.Vector.from_array(x, np.float32)
    y = spb.Vector(m.nrows, np.float32)
    assert spb.spmv_vnec("L", y, A, xv, plan) == ""
    y1 = y.to_array()
    ok, worst = _gate(y1, ptr, col, val, x)
    assert ok, worst
    assert spb.spmv_vnec("L", y, A, xv, plan) == ""
    assert np.array_equal(y.to_array().view(np.uint32), y1.view(np.uint32))
    spb.free_vnec("L", plan, fp32=True)


FP32_SHAPES = {("S", 2), ("S", 3), ("S", 5), ("G", 8), ("L", 5), ("L", 7), ("L", 11), ("C", 5),
               ("C", 7), ("C", 9), ("C", 11), ("D", 7)}


def test_fp32_autotune_times_the_fp32_shapes_and_caches_its_pick():
    """The autotune mode on FP32 matrices times the FP32 plan shapes on the
    matrix itself (slices, pattern slices, L and C with per-tile carries, D
    at width 16); its pick is one of them, cached (a second plan takes the
    same shape; C in 512-thread tiles too), faster than or as fast as the rule's pick within timing
    noise, and within the FP32 gate."""
    for m in (g.stencil27(40, seed=31), g.rmat(16, 16, seed=33), g.banded(200000, 6, seed=34)):
        ptr, col, val, x = _f32(m)
        A = spb.Matrix.from_csr(m.nrows, m.ncols, ptr, col, val, dtype=np.float32)
        p = spb.Plan(A, "autotune")
        info = p.info()
        print(f"FP32 autotune {m.name}: {info.variant}{info.ipt} grid {info.grid}", flush=True)
        assert info.encoding == "fp32" and (info.variant, info.ipt) in FP32_SHAPES
        assert info.lanes in (64, 16, spb.LANES_WIDE, spb.LANES_WIDE | spb.LANES_GATHER)
        q = spb.Plan(A, "autotune")
        assert (q.info().variant, q.info().ipt, q.info().lanes) == (info.variant, info.ipt, info.lanes)
        q.free()
        r = spb.Plan(A, "learned")  # the FP32 rule
        dx = spb.DeviceArray.from_array(x)
        dy = spb.DeviceArray(4 * m.nrows)
        t_auto = min(p.time(dx.ptr, dy.ptr, 20) for _ in range(3))
        t_rule = min(r.time(dx.ptr, dy.ptr, 20) for _ in range(3))
        assert t_auto <= 1.10 * t_rule, (info, r.info(), t_auto, t_rule)
        yv = spb.Vector(m.nrows, np.float32)
        p.spmv(yv, spb.Vector.from_array(x, np.float32))
        ok, worst = _gate(yv.to_array(), ptr, col, val, x)
        assert ok, (info, worst)
        for h in (p, r):
            h.free()
        dx.free()
        dy.free()
        A.free()
