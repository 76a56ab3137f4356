"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle and
the reference's golden fixtures.

Integer structures (merge-path coordinates, ECR maps, D layouts, IRD
decision) must be bit-exact; FP64 results must satisfy, per row,
|y - y_ref| <= 1e-12 * (|A||x|)_i (FP32 hi+lo too); FP32 hi-only is
bounded by 2^-23 * (|A||x|)_i.  Integer-valued known answers must be exact.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from conftest import golden_names, load_golden
from vnec_amd import matgen as g
from vnec_amd import spb

pytestmark = pytest.mark.gpu

TOL = 1e-12
HI_TOL = 2.0 ** -23


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if spb.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    spb.init()


def _mat(m):
    return spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)


def _run_plan(A, x, variant, ipt, enc="fp64", nrhs=1, grid=0, lanes=None):
    if lanes is None:
        lanes = {"D": 8, "H": 4096}.get(variant, 64)
    p = spb.Plan(A, variant, ipt, lanes=lanes, encoding=enc, nrhs=nrhs, grid=grid)
    xv = spb.Vector.from_array(x)
    yv = spb.Vector(A.nrows * nrhs)
    p.spmv(yv, xv)
    y = yv.to_array()
    p.free()
    return y


def _cases():
    out = [("rmat12", g.rmat(12, 16, seed=11)),
           ("stencil16", g.stencil27(16, seed=12)),
           ("powerlaw", g.powerlaw(6000, 1.7, 2, 5000, seed=13)),
           ("banded", g.banded(5000, 40, seed=14)),
           ("blockdiag", g.blockdiag(4000, 64, seed=15))]
    # one dense row spanning many tiles, empty rows around it
    n = 3000
    lens = np.zeros(n, np.int64)
    lens[::3] = 5
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.sort(np.random.default_rng(i).choice(n, 5, replace=False))
                          for i in range(n // 3)]).astype(np.int32)
    dense_row = 1500
    ptr2 = ptr.copy()
    ptr2[dense_row + 1:] += n
    col2 = np.concatenate([col[:ptr[dense_row]], np.arange(n, dtype=np.int32),
                           col[ptr[dense_row]:]])
    out.append(("denserow", g.from_pattern(n, n, ptr2, col2, seed=16)))
    return out


CASES = _cases()


@pytest.mark.parametrize("name,m", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("variant,ipt", [("L", 3), ("L", 7), ("L", 15), ("E", 5), ("E", 11), ("C", 5),
                                         ("C", 15)])
@pytest.mark.parametrize("grid", [0, 8, -1, -2])
def test_spmv_fp64_matches_oracle(name, m, variant, ipt, grid):
    """grid 0: persistent pipelined kernel, 8: persistent with few workgroups
    (long tile runs and carries across many tiles), -1: one workgroup per tile,
    -2: per-tile carries from the strided persistent launch (L and C; E runs
    -1 there).  Variant C is L with each tile's nonzeros stored in column order."""
    if grid == -2 and variant == "E":
        grid = -1
    A = _mat(m)
    print(f"case {name} variant {variant} ipt {ipt} grid {grid}", flush=True)
    y = _run_plan(A, m.x, variant, ipt, grid=grid)
    y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ok, worst = O.check_rows(y, y0, ax, TOL)
    assert ok, worst


@pytest.mark.parametrize("name,m", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("ipt", [1, 3, 7])
def test_slice_variant_matches_oracle(name, m, ipt):
    """Variant S (64-row slices; window or direct columns by plan)."""
    A = _mat(m)
    print(f"case {name} variant S ipt {ipt}", flush=True)
    y = _run_plan(A, m.x, "S", ipt)
    y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ok, worst = O.check_rows(y, y0, ax, TOL)
    assert ok, worst


@pytest.mark.parametrize("name,m", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("ipt,lanes", [(5, 4), (7, 8), (9, 16), (5, 16)])
def test_d_variant_matches_oracle(name, m, ipt, lanes):
    """Variant D (dense groups of `lanes` window positions, wave64 layout).
    Its greedy grouping needs ascending columns in every row (the CSR
    builder sorts them); a matrix with a descending row is refused."""
    A = _mat(m)
    print(f"case {name} variant D ipt {ipt} lanes {lanes}", flush=True)
    rows_sorted = all((np.diff(m.col[m.ptr[i]:m.ptr[i + 1]]) >= 0).all() for i in range(m.nrows))
    if not rows_sorted:
        with pytest.raises(RuntimeError, match="ascending"):
            _run_plan(A, m.x, "D", ipt, lanes=lanes)
        return
    y = _run_plan(A, m.x, "D", ipt, lanes=lanes)
    y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ok, worst = O.check_rows(y, y0, ax, TOL)
    assert ok, worst


@pytest.mark.parametrize("name,m", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("ipt,hot", [(5, 4096), (7, 8192), (5, 8192)])
def test_h_variant_matches_oracle(name, m, ipt, hot):
    """Variant H (persistent merge path, most used x entries in LDS); the
    R-MAT and dense-row cases have hot columns, the others none."""
    A = _mat(m)
    print(f"case {name} variant H ipt {ipt} hot {hot}", flush=True)
    y = _run_plan(A, m.x, "H", ipt, lanes=hot)
    y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ok, worst = O.check_rows(y, y0, ax, TOL)
    assert ok, worst


@pytest.mark.parametrize("variant", ["L", "E", "S"])
@pytest.mark.parametrize("enc", ["hilo", "hi"])
def test_encodings_error_bounds(variant, enc):
    m = g.stencil27(20, seed=21)
    A = _mat(m)
    y = _run_plan(A, m.x, variant, 7, enc)
    y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ok, worst = O.check_rows(y, y0, ax, TOL if enc == "hilo" else HI_TOL)
    assert ok, worst
    if enc == "hi":
        assert worst > 1e-12  # hi-only really drops the low bits


def _spmm_ref(m, X, k):
    Y0 = O.spmm_serial(m.ptr, m.col, m.val, X, k)
    ax = np.empty(m.nrows * k)
    for c in range(k):
        ax[c::k] = O.spmv_abs(m.ptr, m.col, m.val, X[c::k])[0]
    return Y0, ax


@pytest.mark.parametrize("k", [2, 4, 8, 16])
def test_multi_rhs(k):
    m = g.rmat(11, 16, seed=31)
    rng = np.random.default_rng(k)
    X = rng.uniform(-1, 1, m.ncols * k)
    A = _mat(m)
    Y = _run_plan(A, X, "L", 5, nrhs=k)
    Y0, ax = _spmm_ref(m, X, k)
    assert O.check_rows(Y, Y0, ax, TOL)[0]


def _wide_row(n=20000, seed=33):
    """Rows of 0-6 entries and one full row of n: the full row spans 11-26
    tiles of 768-1792 items, beyond the fix-up's 8-tile runs."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 7, n)
    lens[n // 2] = n
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.arange(n) if l == n else np.sort(rng.choice(n, l, replace=False))
                          for l in lens]).astype(np.int32)
    return g.from_pattern(n, n, ptr, col, seed=seed)


@pytest.mark.parametrize("k", [4, 8, 16])
@pytest.mark.parametrize("ipt", [3, 7])
@pytest.mark.parametrize("name", ["widerow", "denserow", "powerlaw"])
def test_multi_rhs_long_rows(name, ipt, k):
    """Rows spanning several tiles, with k right-hand sides: runs the fix-up
    sums per boundary thread (up to 8 tiles) and by the whole wavefront."""
    m = _wide_row() if name == "widerow" else dict(CASES)[name]
    X = np.random.default_rng(100 + k).uniform(-1, 1, m.ncols * k)
    Y = _run_plan(_mat(m), X, "L", ipt, nrhs=k)
    Y0, ax = _spmm_ref(m, X, k)
    ok, worst = O.check_rows(Y, Y0, ax, TOL)
    assert ok, worst


@pytest.mark.parametrize("k", [4, 8, 16])
def test_multi_rhs_unaligned_device_vectors(k):
    """x and y only 8-byte aligned (device pointers from the caller): the
    one-lane-per-nonzero kernel runs instead of the 16-byte grouped one;
    both meet the gate and repeat bit for bit."""
    m = g.rmat(12, 16, seed=32)
    X = np.random.default_rng(k).uniform(-1, 1, m.ncols * k)
    Y0, ax = _spmm_ref(m, X, k)
    p = spb.Plan(_mat(m), "L", 5, nrhs=k)
    out = []
    for off in (0, 8):
        dx = spb.DeviceArray(8 * m.ncols * k + 16)
        dy = spb.DeviceArray(8 * m.nrows * k + 16)
        L = spb.L.load()
        spb.L.check(L.vnec_memcpy_h2d_ext(dx.ptr + off, X.ctypes.data, X.nbytes), "h2d")
        p.spmv_dev(dx.ptr + off, dy.ptr + off)
        Y = np.empty(m.nrows * k)
        spb.L.check(L.vnec_memcpy_d2h_ext(Y.ctypes.data, dy.ptr + off, Y.nbytes), "d2h")
        ok, worst = O.check_rows(Y, Y0, ax, TOL)
        assert ok, (off, worst)
        p.spmv_dev(dx.ptr + off, dy.ptr + off)
        Y2 = np.empty_like(Y)
        spb.L.check(L.vnec_memcpy_d2h_ext(Y2.ctypes.data, dy.ptr + off, Y2.nbytes), "d2h")
        assert np.array_equal(Y, Y2)
        out.append(Y)
        dx.free()
        dy.free()
    p.free()


def test_known_answer_is_exact():
    m = g.rmat(13, 8, seed=41)
    val = np.repeat(np.arange(m.nrows, dtype=np.float64), np.diff(m.ptr))
    x = np.arange(m.ncols, dtype=np.float64)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, val)
    rows = np.repeat(np.arange(m.nrows), np.diff(m.ptr))
    expect = np.bincount(rows, weights=rows * m.col.astype(np.float64), minlength=m.nrows)
    for variant, ipt in (("L", 7), ("E", 7), ("S", 3), ("D", 7), ("H", 5), ("C", 9)):
        assert np.array_equal(_run_plan(A, x, variant, ipt), expect)


def test_empty_rows_and_empty_matrix():
    A = spb.Matrix.from_csr(5, 4, [0, 0, 0, 0, 0, 0], [], [])
    assert np.array_equal(_run_plan(A, np.ones(4), "L", 7), np.zeros(5))
    m = g.from_pattern(6, 6, np.array([0, 0, 2, 2, 2, 3, 3]), np.array([1, 4, 0], np.int32), 3)
    A = _mat(m)
    y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    for variant in ("L", "E", "S"):
        assert np.allclose(_run_plan(A, m.x, variant, 3), y0, rtol=0, atol=1e-15)
    assert np.allclose(_run_plan(A, m.x, "D", 5), y0, rtol=0, atol=1e-15)
    assert np.allclose(_run_plan(A, m.x, "H", 5), y0, rtol=0, atol=1e-15)
    assert np.allclose(_run_plan(A, m.x, "C", 5), y0, rtol=0, atol=1e-15)
    A = spb.Matrix.from_csr(5, 4, [0, 0, 0, 0, 0, 0], [], [])
    assert np.array_equal(_run_plan(A, np.ones(4), "C", 7), np.zeros(5))
    assert np.array_equal(_run_plan(A, np.ones(4), "S", 1), np.zeros(5))
    assert np.array_equal(_run_plan(A, np.ones(4), "D", 7), np.zeros(5))
    assert np.array_equal(_run_plan(A, np.ones(4), "H", 7), np.zeros(5))


# -------------------------------------------------- builder structures
@pytest.mark.parametrize("name", golden_names())
def test_partition_bit_exact_with_reference(name):
    d = load_golden(name)
    A = spb.Matrix.from_csr(int(d["nrows"]), int(d["ncols"]), d["ptr"], d["col"], d["val"])
    co = spb.partition(A, int(d["P"]))
    assert np.array_equal(np.concatenate([co[:-1], co[1:]], 1), d["coords"])


@pytest.mark.parametrize("P", [1, 7, 36, 500, 4097])
@pytest.mark.parametrize("snap", [False, True])
def test_partition_bit_exact_with_oracle(P, snap):
    m = g.powerlaw(20000, 2.2, 2, 3000, seed=P)
    A = _mat(m)
    co = spb.partition(A, P, snap)
    s, e = O.partition(m.ptr, P, snap)
    assert np.array_equal(co[:-1], s) and np.array_equal(co[1:], e)


@pytest.mark.parametrize("name", golden_names())
def test_ecr_bit_exact_with_reference(name):
    d = load_golden(name)
    A = spb.Matrix.from_csr(int(d["nrows"]), int(d["ncols"]), d["ptr"], d["col"], d["val"])
    co = spb.partition(A, int(d["P"]))
    for method in (1, 2):
        if method == 2 and np.diff(co[:, 1]).max() > 4096:
            continue
        e, nec, off, win = spb.ecr(A, co, method)
        assert np.array_equal(e, d["ecr"]) and np.array_equal(nec, d["nec"])
        assert np.array_equal(win, d["win"])


@pytest.mark.parametrize("ipt", [3, 7, 15])
def test_ecr_tiles_bit_exact_with_oracle(ipt):
    m = g.rmat(14, 16, seed=ipt)
    A = _mat(m)
    P = -(-(m.nrows + m.nnz) // (256 * ipt))
    co = spb.partition(A, P)
    e, nec, off, win = spb.ecr(A, co, 2)
    ref = O.ecr(m.col, co[:-1], co[1:])
    assert np.array_equal(e, ref.ecr) and np.array_equal(nec, ref.nec)
    assert np.array_equal(off, ref.win_off) and np.array_equal(win, ref.win)


@pytest.mark.parametrize("name", golden_names())
def test_d_layout_and_ird_bit_exact_with_reference(name):
    d = load_golden(name)
    A = spb.Matrix.from_csr(int(d["nrows"]), int(d["ncols"]), d["ptr"], d["col"], d["val"])
    vrp, cs, va = spb.d_layout(A, int(d["P"]), int(d["L"]))
    assert np.array_equal(vrp, d["v_row_ptr"]) and np.array_equal(cs, d["col_start"])
    assert np.array_equal(va, d["val_align"])
    assert spb.ird_vnec(A) == int(d["ird"])


@pytest.mark.parametrize("name", golden_names())
def test_s_order_bit_exact_with_reference(name):
    """GPU S task order and reordered matrix at the reference's P = 36,
    L = 8, cut 32 against the reference build's own structures where the
    fixture carries them; on the fixtures where the reference's length
    buckets overflow (equal-length rows, spmv_vnec_S_for_mxv.c:442-524, so
    its structures are undefined and were not exported) against the
    oracle's restatement at the same parameters instead."""
    d = load_golden(name)
    n = int(d["nrows"])
    A = spb.Matrix.from_csr(n, int(d["ncols"]), d["ptr"], d["col"], d["val"])
    r = spb.s_order(A, 36, 8, 32)
    if "s_tasks" not in d:
        o = O.s_build(d["ptr"], d["col"], d["val"], 36, 8, 32)
        assert np.array_equal(r["tasks"], o.tasks) and np.array_equal(r["spvv_len"], o.spvv_len)
        assert np.array_equal(r["row_begin"], o.row_begin) and np.array_equal(r["row_end"], o.row_end)
        assert np.array_equal(r["s_col"], o.s_col) and np.array_equal(r["s_val"], o.s_val)
        assert np.array_equal(r["task_start"], o.task_start) and np.array_equal(r["task_end"], o.task_end)
        return
    assert np.array_equal(r["tasks"], d["s_tasks"])
    assert np.array_equal(r["spvv_len"], d["s_spvv_len"])
    assert np.array_equal(r["row_begin"], d["s_row_begin"])
    assert np.array_equal(r["row_end"], d["s_row_end"])
    assert np.array_equal(r["s_col"], d["s_s_col"])
    assert np.array_equal(r["s_val"], d["s_s_val"])
    assert np.array_equal(np.c_[r["task_start"], r["task_end"]], d["s_task_se"])


@pytest.mark.parametrize("P,L_,cut", [(1, 8, 32), (7, 4, 8), (36, 8, 32), (300, 16, 64), (2000, 2, 3)])
def test_s_order_bit_exact_with_oracle(P, L_, cut):
    m = g.powerlaw(12000, 1.9, 1, 400, seed=P + L_)
    A = _mat(m)
    r = spb.s_order(A, P, L_, cut)
    o = O.s_build(m.ptr, m.col, m.val, P, L_, cut)
    assert np.array_equal(r["tasks"], o.tasks) and np.array_equal(r["spvv_len"], o.spvv_len)
    assert np.array_equal(r["row_begin"], o.row_begin) and np.array_equal(r["row_end"], o.row_end)
    assert np.array_equal(r["s_col"], o.s_col) and np.array_equal(r["s_val"], o.s_val)
    assert np.array_equal(r["task_start"], o.task_start) and np.array_equal(r["task_end"], o.task_end)


@pytest.mark.parametrize("P,L_", [(1, 1), (7, 4), (64, 64), (300, 8)])
def test_d_layout_bit_exact_with_oracle(P, L_):
    m = g.banded(4000, 9, seed=P + L_)
    A = _mat(m)
    vrp, cs, va = spb.d_layout(A, P, L_)
    ref = O.d_build(m.ptr, m.col, m.val, P, L_)
    assert np.array_equal(vrp, ref.v_row_ptr) and np.array_equal(cs, ref.col_start)
    assert np.array_equal(va, ref.val_align)


# ------------------------------------------------- reference entry points
@pytest.mark.parametrize("name", ["rmat10", "stencil10", "banded3000_hb5", "rmat11_kat"])
def test_reference_api_build_and_spmv(name):
    d = load_golden(name)
    n = int(d["nrows"])
    A = spb.Matrix.from_csr(n, int(d["ncols"]), d["ptr"], d["col"], d["val"])
    x = spb.Vector.from_array(d["x"])
    y0 = O.spmv_serial(d["ptr"], d["col"], d["val"], d["x"])
    ax, _ = O.spmv_abs(d["ptr"], d["col"], d["val"], d["x"])
    for kind in ("L", "D", "S"):
        plan = spb.build_vnec(kind, A)
        y = spb.Vector(n)
        assert spb.spmv_vnec(kind, y, A, x, plan) == ""
        yy = y.to_array()
        assert O.check_rows(yy, y0, ax, TOL)[0], kind
        if "y_kat" in d:
            assert np.array_equal(yy, d["y_kat"])
        p = plan.contents
        if kind == "L":
            co = np.array([[p.thread_coord_start[t].x, p.thread_coord_start[t].y,
                            p.thread_coord_end[t].x, p.thread_coord_end[t].y] for t in range(36)])
            assert np.array_equal(co, d["l_coords"])
        if kind == "D":
            nz = int(d["ptr"][-1])
            assert np.array_equal(np.ctypeslib.as_array(p.ecr_indices, (nz,)), d["ecr"])
            assert np.array_equal(np.ctypeslib.as_array(p.v_row_ptr, (n + 1,)), d["v_row_ptr"])
            G = len(d["col_start"])
            if G:
                assert np.array_equal(np.ctypeslib.as_array(p.col_start, (G,)), d["col_start"])
                # val_align: the reference's G * 8 filled entries (it allocates nnz * 8)
                assert np.array_equal(np.ctypeslib.as_array(p.val_align, (G * 8,)),
                                      d["val_align"][:G * 8])
            nec = np.ctypeslib.as_array(p.NEC_NUM, (36,))
            assert np.array_equal(nec, d["nec"])
        if kind == "S" and "s_tasks" in d:
            nz = int(d["ptr"][-1])
            assert np.array_equal(np.ctypeslib.as_array(p.NEC_NUM, (36,)), d["s_nec"])
            assert np.array_equal(np.ctypeslib.as_array(p.spvv8_len, (36,)), d["s_spvv_len"])
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
        spb.free_vnec(kind, plan)
    # SpB_mxv with no plan builds and caches one
    y = spb.Vector(n)
    spb.mxv(y, A, x)
    assert O.check_rows(y.to_array(), y0, ax, TOL)[0]


def test_resident_vectors_and_repeated_calls():
    m = g.stencil27(12, seed=51)
    A = _mat(m)
    p = spb.Plan(A, "E", 7)
    x = spb.Vector.from_array(m.x)
    y = spb.Vector(m.nrows)
    x.resident(True)
    y.resident(True)
    y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    for _ in range(3):
        p.spmv(y, x)
    assert O.check_rows(y.to_array(), y0, ax, TOL)[0]
    x.set(0, 0.0)  # host write marks the device copy stale
    xm = m.x.copy()
    xm[0] = 0.0
    p.spmv(y, x)
    assert O.check_rows(y.to_array(), O.spmv_serial(m.ptr, m.col, m.val, xm), ax, TOL)[0]


def test_graph_timing_path_gives_same_result():
    m = g.rmat(12, 16, seed=61)
    A = _mat(m)
    p = spb.Plan(A, "L", 7)
    dx = spb.DeviceArray.from_array(m.x)
    dy = spb.DeviceArray(8 * m.nrows)
    ms = p.time(dx.ptr, dy.ptr, 30, graph=True)
    assert ms > 0
    y = dy.to_array(np.float64, m.nrows)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    assert O.check_rows(y, O.spmv_serial(m.ptr, m.col, m.val, m.x), ax, TOL)[0]


@pytest.mark.parametrize("variant,ipt", [("L", 7), ("E", 7), ("S", 3), ("D", 7), ("H", 7)])
def test_full_size_stencil_per_row_gate(variant, ipt):
    m = g.stencil27(128, seed=1)
    A = _mat(m)
    y = _run_plan(A, m.x, variant, ipt)
    ax, yl = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    assert O.check_rows(y, yl, ax, TOL)[0]


def test_repeated_columns_are_summed_by_every_variant():
    """Duplicates are kept by the CSR builder (SpB_Matrix_build_FP64), so a
    row may name a column twice; every kernel must add both entries."""
    rng = np.random.default_rng(5)
    n = 900
    lens = rng.integers(0, 12, n)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.sort(np.concatenate([c, c[:2]])) if len(c) > 2 else np.sort(c)
                          for c in (rng.choice(n, k, replace=False) for k in lens)]).astype(np.int32)
    lens2 = np.array([k + 2 if k > 2 else k for k in lens])
    ptr = np.concatenate([[0], np.cumsum(lens2)]).astype(np.int64)
    m = g.from_pattern(n, n, ptr, col, seed=6)
    A = _mat(m)
    y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    for variant, ipt in (("L", 7), ("E", 5), ("S", 3), ("D", 5), ("D", 9), ("H", 7), ("C", 7)):
        ok, worst = O.check_rows(_run_plan(A, m.x, variant, ipt), y0, ax, TOL)
        assert ok, (variant, worst)


def test_column_sorted_tiles_rule_and_limits():
    """Variant C: learned plans switch from L to C on skewed column use (at
    most 0.8 distinct columns per nonzero over 16384-item windows, at least
    2^20 nonzeros), not on uniform columns nor small matrices; C keeps the
    per-row gate; it refuses what it cannot run."""
    m = g.rmat(17, 16, seed=71)
    ints = spb.feature_ints(_mat(m))
    assert ints[30] <= 0.8 * ints[2] and ints[2] >= 1 << 20  # VF_NEC_LARGE, VF_NNZ
    A = _mat(m)
    p = spb.Plan(A, "learned")
    info = p.info()
    if info.variant in ("L", "C"):
        assert info.variant == "C"
    xv = spb.Vector.from_array(m.x)
    yv = spb.Vector(m.nrows)
    p.spmv(yv, xv)
    ax, yl = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    assert O.check_rows(yv.to_array(), yl, ax, TOL)[0]
    p.free()
    u = g.uniform(1 << 17, 16, seed=72)
    U = _mat(u)
    assert spb.feature_ints(U)[30] > 0.8 * spb.feature_ints(U)[2]
    assert spb.Plan(U, "learned").info().variant != "C"
    s = _mat(g.rmat(14, 16, seed=73))  # skewed but small
    assert spb.Plan(s, "learned").info().variant != "C"
    with pytest.raises(RuntimeError, match="variant C stores FP64 values"):
        spb.Plan(A, "C", 7, encoding="hi")
    with pytest.raises(RuntimeError, match="variant C takes one right-hand side"):
        spb.Plan(A, "C", 7, nrhs=4)
    with pytest.raises(RuntimeError, match="ipt 8"):
        spb.Plan(A, "C", 8)


@pytest.mark.parametrize("ipt", [3, 9, 15])
def test_column_sorted_tiles_wide_matrix(ipt):
    """Variant C keeps each nonzero's tile offset beside its column, packed
    in one word while both fit 32 bits, else in a separate array: 2^23
    columns force the separate array, 2^12 columns the packed form."""
    rng = np.random.default_rng(80 + ipt)
    for ncols in (1 << 23, 1 << 12):
        n = 30000
        lens = rng.integers(0, 20, n)
        ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        col = np.concatenate([np.sort(rng.choice(ncols, k, replace=False)) for k in lens]).astype(np.int32)
        m = g.from_pattern(n, ncols, ptr, col, seed=ipt)
        y = _run_plan(_mat(m), m.x, "C", ipt)
        ax, yl = O.spmv_abs(m.ptr, m.col, m.val, m.x)
        ok, worst = O.check_rows(y, yl, ax, TOL)
        assert ok, (ncols, worst)


@pytest.mark.parametrize("variant,ipt,enc", [("L", 3, "fp64"), ("L", 11, "hilo"), ("L", 7, "hi"), ("C", 9, "fp64")])
def test_strided_launch_is_bit_identical_to_one_workgroup_per_tile(variant, ipt, enc):
    """grid -2 runs the tiles of grid -1 (same tiles, same per-tile carries,
    same fix-up) from a persistent launch: the results agree bit for bit."""
    m = g.rmat(15, 16, seed=91)
    A = _mat(m)
    p1 = spb.Plan(A, variant, ipt, encoding=enc, grid=-1)
    p2 = spb.Plan(A, variant, ipt, encoding=enc, grid=-2)
    assert p2.info().grid == -2
    y1 = _run_plan(A, m.x, variant, ipt, enc, grid=-1)
    y2 = _run_plan(A, m.x, variant, ipt, enc, grid=-2)
    assert np.array_equal(y1, y2)
    p1.free()
    p2.free()
    with pytest.raises(RuntimeError, match="grid -2"):
        spb.Plan(A, "E", 5, grid=-2)
    with pytest.raises(RuntimeError, match="grid -3"):
        spb.Plan(A, "L", 5, grid=-3)


@pytest.mark.parametrize("k", [1, 2, 4, 8, 16])
def test_multi_rhs_column_major(k):
    """Column-major right-hand sides (k vectors one after another) give the
    row-interleaved plan's results bit for bit, and meet the gate against k
    serial SpMVs; 1037 rows and columns leave a partial 256-entry block."""
    m = g.powerlaw(1037, 2.0, 2, 200, seed=95)
    assert m.nrows == m.ncols == 1037
    n = m.nrows
    rng = np.random.default_rng(200 + k)
    Xc = rng.uniform(-1, 1, (k, m.ncols))
    p = spb.Plan(_mat(m), "L", 5, nrhs=k)
    L = spb.L.load()
    dx = spb.DeviceArray.from_array(Xc.ravel())
    dy = spb.DeviceArray(8 * n * k)
    p.spmm_colmajor_dev(dx.ptr, dy.ptr)
    Yc = dy.to_array(np.float64, n * k).reshape(k, n)
    xi = spb.DeviceArray.from_array(np.ascontiguousarray(Xc.T).ravel())
    yi = spb.DeviceArray(8 * n * k)
    p.spmv_dev(xi.ptr, yi.ptr)
    Yi = yi.to_array(np.float64, n * k).reshape(n, k)
    assert np.array_equal(Yc, Yi.T)
    for c in range(k):
        ax, yl = O.spmv_abs(m.ptr, m.col, m.val, Xc[c])
        assert O.check_rows(Yc[c], yl, ax, TOL)[0]
    for a in (dx, dy, xi, yi):
        a.free()
    p.free()


@pytest.mark.parametrize("variant,ipt,grid", [("L", 7, 0), ("L", 7, -1), ("C", 9, 0),
                                              ("S", 3, 0), ("D", 7, 0), ("G", 8, 0)])
def test_time_kernels_leaves_y_right(variant, ipt, grid):
    """The kernel timer repeats each kernel back to back; y must still come
    back as A·x, and both means are non-negative (the SpMV kernel's positive)."""
    m = g.stencil27(24, seed=31) if variant == "G" else g.rmat(13, 16, seed=31)
    A = _mat(m)
    p = spb.Plan(A, variant, ipt, lanes=8 if variant == "D" else 64, grid=grid)
    dx = spb.DeviceArray.from_array(m.x)
    dy = spb.DeviceArray(8 * m.nrows)
    k_ms, f_ms = p.time_kernels(dx.ptr, dy.ptr, 25)
    y = dy.to_array(np.float64, m.nrows)
    dx.free()
    dy.free()
    p.free()
    assert k_ms > 0 and f_ms >= 0
    ax, yl = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ok, worst = O.check_rows(y, yl, ax, TOL)
    assert ok, worst


_WIDE_BATCH = []


def _wide_batch():
    """64 members of 16k-20k rows: more than 2^20 columns in all, so a
    column and a 12-bit tile offset no longer share 32 bits, but every tile
    spans at most two members' columns."""
    if not _WIDE_BATCH:
        _WIDE_BATCH.append(g.batch(64, 16_000, 20_000, 4.0, 16.0, seed=21))
    return _WIDE_BATCH[0]


@pytest.mark.parametrize("ipt,grid", [(9, 512), (11, 512), (11, -1), (9, -2)])
def test_c_tile_relative_columns(ipt, grid):
    """Variant C packs each offset above a column taken relative to its
    tile's smallest column when absolute columns are too wide: the plan
    holds values, packed words and one base per tile (no offset array), and
    y passes the row gate."""
    m = _wide_batch()
    assert m.ncols > 1 << 20
    A = spb.Matrix.from_batch(m.row_off, m.col_off, m.ptr, m.col, m.val)
    pl = spb.Plan(A, "L", ipt, grid=grid)
    pc = spb.Plan(A, "C", ipt, grid=grid)
    il, ic = pl.info(), pc.info()
    pl.free()
    assert ic.ntiles == il.ntiles and ic.grid == il.grid
    assert ic.dev_bytes - il.dev_bytes == 12 * m.nnz + 4 * ic.ntiles
    print(f"wide batch nnz {m.nnz} ncols {m.ncols} C{ipt} grid {ic.grid}", flush=True)
    dx = spb.DeviceArray.from_array(m.x)
    dy = spb.DeviceArray(8 * m.nrows)
    pc.spmv_dev(dx.ptr, dy.ptr)
    y = dy.to_array(np.float64, m.nrows)
    dx.free()
    dy.free()
    pc.free()
    col = m.col + np.repeat(m.col_off[:-1], np.diff(m.ptr[m.row_off])).astype(np.int32)
    ax, yl = O.spmv_abs(m.ptr, col, m.val, m.x)
    ok, worst = O.check_rows(y, yl, ax, TOL)
    assert ok, worst


@pytest.mark.parametrize("name", ["wide_batch", "rmat12", "powerlaw"])
def test_tile_colspan_matches_numpy(name):
    """vnec_hip_tile_colspan: each tile's smallest column and the widest
    span, exactly as numpy finds them (empty tiles report base 0)."""
    import ctypes as C
    from vnec_amd import _lib as L
    if name == "wide_batch":
        m = _wide_batch()
        col = m.col + np.repeat(m.col_off[:-1], np.diff(m.ptr[m.row_off])).astype(np.int32)
    else:
        m = dict(_cases())[name]
        col = m.col
    A = _mat(m) if name != "wide_batch" else spb.Matrix.from_batch(m.row_off, m.col_off, m.ptr, m.col, m.val)
    P = -(-(m.nrows + m.nnz) // (256 * 11))
    co = spb.partition(A, P)
    dco = spb.DeviceArray.from_array(np.ascontiguousarray(co, dtype=np.int32))
    dcol = spb.DeviceArray.from_array(np.ascontiguousarray(col, dtype=np.int32))
    dmin = spb.DeviceArray(4 * P)
    span = C.c_int32(0)
    rc = L.load().vnec_hip_tile_colspan(dco.ptr, P, dcol.ptr, dmin.ptr, C.byref(span), None)
    assert rc == 0
    got = dmin.to_array(np.int32, P)
    for d in (dco, dcol, dmin):
        d.free()
    y = co[:, 1].astype(np.int64)
    ref_min = np.zeros(P, np.int64)
    ref_span = 0
    for t in range(P):
        seg = col[y[t]:y[t + 1]]
        if len(seg):
            ref_min[t] = seg.min()
            ref_span = max(ref_span, int(seg.max()) - int(seg.min()) + 1)
    assert np.array_equal(got, ref_min) and span.value == ref_span


def _ird_large_cases():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "ird_large.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", _ird_large_cases(), ids=lambda c: c["name"])
def test_ird_above_2_24_nonzeros_bit_exact_with_reference(case):
    """IRD on the GPU past 2^24 nonzeros (ird_calc.c:103-104 rounds nnz and
    G * L to FP32 there): the GPU group count equals the reference's, the
    ratio is the reference's formula over it bit for bit and prints as the
    reference printed it, and IRD_VNEC decides as the reference did."""
    m = getattr(g, case["gen"])(*case["args"], seed=case["seed"])
    assert (m.nrows, m.nnz) == (case["nrows"], case["nnz"]) and m.nnz > 1 << 24
    print(f"IRD {case['name']}: nrows {m.nrows} nnz {m.nnz}", flush=True)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    G = C.c_int64()
    spb.L.check(spb.L.load().vnec_d_layout_ext(A.h, 36, 8, None, None, None, C.byref(G)), "d layout")
    assert G.value == case["groups"]
    dec, r = spb.ird_ext(A)
    want = np.float32(case["nnz"]) / np.float32(case["groups"] * 8)
    assert np.float32(r).tobytes() == want.tobytes()
    assert "%f" % r == case["ird_printed"]
    assert dec == case["decision"] == spb.ird_vnec(A)
    A.free()


@pytest.mark.parametrize("fp32", [False, True])
def test_reference_d_entry_runs_slice_layouts_on_dense_column_runs(fp32):
    """IRD sends matrices of dense column runs to D; the D entry runs them
    with pattern slices (banded) or 64-row slices (block-diagonal), which
    stream about twice as fast as D blocks there, and its public fields stay
    the reference's (checked against the fixtures elsewhere)."""
    dt = np.float32 if fp32 else np.float64
    for m, want in ((g.banded(1 << 17, 5, seed=81), "G"), (g.blockdiag(1 << 16, 16, seed=82), "S")):
        print(f"D entry: {m.name} fp32={fp32}", flush=True)
        A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val.astype(dt), dtype=dt)
        assert spb.ird_vnec(A) == 0  # SpB_VNEC_D: the reference would run D blocks
        plan = spb.build_vnec("D", A)
        assert spb.reference_plan_info(plan).variant == want
        x = spb.Vector.from_array(m.x.astype(dt), dt)
        y = spb.Vector(m.nrows, dt)
        assert spb.spmv_vnec("D", y, A, x, plan) == ""
        if fp32:
            v32, x32 = m.val.astype(np.float32), m.x.astype(np.float32)
            ax, _ = O.spmv_abs(m.ptr, m.col, v32.astype(np.float64), x32.astype(np.float64))
            assert O.check_rows_f32(y.to_array(), O.spmv_serial_f32(m.ptr, m.col, v32, x32),
                                    m.ptr, ax)[0]
        else:
            ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
            assert O.check_rows(y.to_array(), O.spmv_serial(m.ptr, m.col, m.val, m.x), ax, TOL)[0]
        spb.free_vnec("D", plan, fp32=fp32)
        A.free()


@pytest.mark.parametrize("variant,ipt,grid", [("L", 7, 0), ("L", 11, -1), ("C", 9, 0), ("C", 11, -2),
                                               ("C", 7, -1)])
def test_relabelled_columns_give_the_plain_plans_bits(variant, ipt, grid):
    """Columns relabelled by descending use (x permuted before the SpMV)
    change which x entries share cache lines, not the sums: every row is
    summed in its CSR order, so y equals the plain L plan's bit for bit and
    passes the row gate; the plan refuses other variants."""
    m = g.skew(14, 16, 2, 5000, seed=97)  # R-MAT with two long dense rows
    A = _mat(m)
    y_plain = _run_plan(A, m.x, "L", ipt, grid=grid)
    y = _run_plan(A, m.x, variant, ipt, grid=grid, lanes=1)
    assert np.array_equal(y, y_plain)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    assert O.check_rows(y, O.spmv_serial(m.ptr, m.col, m.val, m.x), ax, TOL)[0]
    p = spb.Plan(A, variant, ipt, lanes=1, grid=grid)
    bm, bs = p.format_bytes()
    assert bs == bm + 20 * m.ncols  # the x permutation's bytes
    p.free()
    with pytest.raises(RuntimeError, match="relabelled columns"):
        spb.Plan(A, "S", 3, lanes=1)
    with pytest.raises(RuntimeError, match="relabelled columns"):
        spb.Plan(A, "L", 7, lanes=1, encoding="hilo")
    A.free()


@pytest.mark.parametrize("name", ["rmat14", "skew14", "holes"])
def test_wide_column_sorted_tiles(name):
    """Variant C with 512-thread tiles (512 ipt items, each sorted by column
    over the whole tile), persistent and with per-tile carries: every ipt
    passes the row gate, long rows spanning
    many tiles and empty rows included; relabelled columns on top give the
    same bits (rows are still summed in CSR order within the same tiles);
    the plan refuses the combinations it has no kernel for."""
    if name == "rmat14":
        m = g.rmat(14, 16, seed=98)
    elif name == "skew14":
        m = g.skew(14, 16, 2, 12000, seed=99)  # two rows spanning several wide tiles
    else:
        m = g.uniform(30000, 7, seed=100)
        lens = np.diff(m.ptr)
        keep = np.repeat(np.arange(m.nrows), lens) % 3 != 0  # every third row empty
        lens[::3] = 0
        ptr = np.concatenate([[0], np.cumsum(lens)])
        m = g.from_pattern(m.nrows, m.ncols, ptr, m.col[keep], seed=101)
    A = _mat(m)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ref = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    for ipt in (3, 5, 7, 9, 11):
        y = _run_plan(A, m.x, "C", ipt, lanes=1024)
        assert O.check_rows(y, ref, ax, TOL)[0], ipt
        # per-tile carries (the strided launch of 512-thread tiles)
        assert O.check_rows(_run_plan(A, m.x, "C", ipt, lanes=1024, grid=-1), ref, ax, TOL)[0], ipt
        if ipt == 9:
            assert np.array_equal(_run_plan(A, m.x, "C", ipt, lanes=1024 | 1), y)
    p = spb.Plan(A, "C", 5, lanes=1024)
    info = p.info()
    assert info.ntiles == -(-(m.nrows + m.nnz) // (512 * 5))  # VNEC_WIDE_BLOCK threads per tile
    p.free()
    p = spb.Plan(A, "C", 7, lanes=1024, grid=-1)
    assert p.info().grid == -2  # wide tiles with per-tile carries run the strided launch
    p.free()
    # gathers issued together (256- and 512-thread tiles): the same tiles and
    # sums as the plan without the flag, so the same bits
    for ipt in (3, 5, 7):
        for lanes in (64, 1024):
            y0 = _run_plan(A, m.x, "C", ipt, lanes=lanes, grid=-2)
            flags = (lanes & spb.LANES_WIDE) | spb.LANES_GATHER
            y1 = _run_plan(A, m.x, "C", ipt, lanes=flags, grid=-1)
            assert np.array_equal(y0, y1), (ipt, lanes)
    for bad in [dict(variant="C", ipt=9, lanes=spb.LANES_GATHER, grid=-1),
                dict(variant="C", ipt=7, lanes=spb.LANES_GATHER, grid=0),
                dict(variant="L", ipt=7, lanes=spb.LANES_GATHER, grid=-1)]:
        with pytest.raises(RuntimeError, match="gathers issued together"):
            spb.Plan(A, bad.pop("variant"), bad.pop("ipt"), **bad)
    for bad in [dict(variant="L", ipt=7, lanes=1024), dict(variant="C", ipt=13, lanes=1024),
                dict(variant="C", ipt=9, lanes=1024, nrhs=4), dict(variant="C", ipt=9, lanes=3)]:
        with pytest.raises(RuntimeError, match="wide tiles|lanes"):
            spb.Plan(A, bad.pop("variant"), bad.pop("ipt"), **bad)
    A.free()


@pytest.mark.parametrize("lanes", [spb.LANES_GATHER, spb.LANES_WIDE | spb.LANES_GATHER])
def test_gathered_tiles_on_rows_without_nonzeros(lanes):
    """Column-sorted tiles with gathers issued together on a matrix whose
    tiles hold no nonzero (every row empty): the kernel skips the gathers of
    such tiles and y is all zeros; a learned plan on the same matrix gives
    the same."""
    for ncols in (5,):
        n = 7000
        A = spb.Matrix.from_csr(n, ncols, np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0))
        for p in (spb.Plan(A, "C", 7, lanes=lanes, grid=-1), spb.Plan(A, "learned")):
            xv = spb.Vector.from_array(np.ones(ncols))
            yv = spb.Vector.from_array(np.full(n, 3.0))
            p.spmv(yv, xv)
            assert np.array_equal(yv.to_array(), np.zeros(n))
            p.free()
        A.free()
