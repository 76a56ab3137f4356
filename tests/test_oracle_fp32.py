"""Pins the FP32 side of the CPU oracle against the reference's FP32 entry
points (fixtures `<name>_f32.npz`, written by tests/golden/make_golden.py
from the reference compiled under oracle/_ref) and against what FP32
arithmetic fixes.

  orc_spmv_serial_f32        the harness's SpB_mxv_FP32_serial
                             (test_spmv_vnec.hpp:34-54): bit-exact against a
                             scalar float32 loop in numpy (product rounded,
                             then sum rounded, in row order); exact KAT
                             y_i = i * sum_j j where every partial sum stays
                             below 2^24; the reference's FP32 kernel outputs
                             within the FP32 row gate
  d_build / s_build at L=16  Build_VNEC_D_FP32 / Build_VNEC_S_FP32 layouts
  ird at L=16, 0.437         IRD_VNEC(A, SpB_FP32) decision
A planted mistake (FP64 accumulation, one dropped entry) must fail.
"""
import numpy as np
import pytest

import oracle as O


def _f32_loop(ptr, col, val, x):
    val = np.asarray(val, np.float32)
    x = np.asarray(x, np.float32)
    y = np.zeros(len(ptr) - 1, np.float32)
    for i in range(len(ptr) - 1):
        t = np.float32(0.0)
        for j in range(int(ptr[i]), int(ptr[i + 1])):
            p = np.float32(val[j] * x[col[j]])
            t = np.float32(t + p)
        y[i] = t
    return y


def _rand(seed, nrows=60, ncols=50, maxlen=40):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, maxlen, nrows)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(ncols, l, replace=True)) for l in lens]).astype(np.int32)
    val = rng.uniform(-1, 1, len(col)).astype(np.float32)
    x = rng.uniform(-1, 1, ncols).astype(np.float32)
    return ptr, col, val, x


@pytest.mark.parametrize("seed", range(4))
def test_serial_f32_is_the_float32_loop_bit_for_bit(seed):
    ptr, col, val, x = _rand(seed)
    y = O.spmv_serial_f32(ptr, col, val, x)
    assert np.array_equal(y.view(np.uint32), _f32_loop(ptr, col, val, x).view(np.uint32))


def test_serial_f32_differs_from_fp64_accumulation():
    """The planted mistake "accumulate in FP64, round once" is visible."""
    ptr, col, val, x = _rand(11, nrows=200, maxlen=200)
    y = O.spmv_serial_f32(ptr, col, val, x)
    y64 = O.spmv_serial(ptr, col, val.astype(np.float64), x.astype(np.float64)).astype(np.float32)
    assert not np.array_equal(y, y64)


def test_serial_f32_exact_on_small_kat(golden32):
    ptr, col = golden32["ptr"], golden32["col"]
    if "y_kat" not in golden32:
        pytest.skip("not a known-answer fixture")
    y_kat = golden32["y_kat"]
    # exact only while every partial sum is an integer below 2^24
    rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    csum = np.zeros(len(col))
    for i in range(len(ptr) - 1):
        b, e = ptr[i], ptr[i + 1]
        csum[b:e] = np.cumsum(i * col[b:e].astype(np.float64))
    y = O.spmv_serial_f32(ptr, col, golden32["val"], golden32["x"])
    exact_rows = np.ones(len(ptr) - 1, bool)
    np.logical_and.at(exact_rows, rows, csum < 2 ** 24)
    assert exact_rows.sum() > 0
    assert np.array_equal(y[exact_rows].astype(np.float64), y_kat[exact_rows])


def test_reference_fp32_outputs_within_gate(golden32):
    ptr, col, val, x = golden32["ptr"], golden32["col"], golden32["val"], golden32["x"]
    y0 = O.spmv_serial_f32(ptr, col, val, x)
    ax, _ = O.spmv_abs(ptr, col, val.astype(np.float64), x.astype(np.float64))
    seen = 0
    for k in ("y_L", "y_D", "y_S"):
        if k in golden32:
            seen += 1
            ok, worst = O.check_rows_f32(golden32[k], y0, ptr, ax)
            assert ok, (golden32["name"], k, worst)
    if len(ptr) > 100:
        assert seen >= 1


def test_fp32_gate_catches_a_dropped_entry(golden32):
    ptr, col, val, x = golden32["ptr"], golden32["col"], golden32["val"], golden32["x"]
    if "y_L" not in golden32:
        pytest.skip("no reference L output")
    y0 = O.spmv_serial_f32(ptr, col, val, x)
    ax, _ = O.spmv_abs(ptr, col, val.astype(np.float64), x.astype(np.float64))
    bad = golden32["y_L"].copy()
    j = int(np.argmax(np.abs(val) * np.abs(x[col])))
    r = int(np.searchsorted(ptr, j, side="right") - 1)
    bad[r] -= val[j] * x[col[j]]
    assert not O.check_rows_f32(bad, y0, ptr, ax)[0]


def test_d_layout_l16_matches_reference(golden32):
    ptr, col, val = golden32["ptr"], golden32["col"], golden32["val"]
    P, L = int(golden32["P"]), int(golden32["L"])
    assert L == 16
    s, e = O.partition(ptr, P)
    assert np.array_equal(np.concatenate([s, e], 1), golden32["coords"])
    ec = O.ecr(col, s, e)
    assert np.array_equal(ec.ecr, golden32["ecr"])
    assert np.array_equal(ec.nec, golden32["nec"])
    assert np.array_equal(ec.win, golden32["win"])
    d = O.d_build(ptr, col, val.astype(np.float64), P, L)
    assert np.array_equal(d.v_row_ptr, golden32["v_row_ptr"])
    assert np.array_equal(d.col_start, golden32["col_start"])
    assert np.array_equal(d.val_align.astype(np.float32), golden32["val_align"])


def test_s_layout_l16_matches_reference(golden32):
    if "s_tasks" not in golden32:
        pytest.skip("reference S builder unsafe on this input")
    ptr, col, val = golden32["ptr"], golden32["col"], golden32["val"]
    sp = O.s_build(ptr, col, val.astype(np.float64), int(golden32["P"]), int(golden32["L"]))
    assert np.array_equal(sp.ecr.nec, golden32["s_nec"])
    assert np.array_equal(sp.ecr.win, golden32["s_win"])
    assert np.array_equal(sp.spvv_len, golden32["s_spvv_len"])
    assert np.array_equal(sp.tasks, golden32["s_tasks"])
    assert np.array_equal(sp.row_begin, golden32["s_row_begin"])
    assert np.array_equal(sp.row_end, golden32["s_row_end"])
    assert np.array_equal(sp.s_col, golden32["s_s_col"])
    assert np.array_equal(sp.s_val.astype(np.float32), golden32["s_s_val"])
    assert np.array_equal(np.stack([sp.task_start, sp.task_end], 1), golden32["s_task_se"])


def test_ird_fp32_matches_reference(golden32):
    dec, r = O.ird(golden32["ptr"], golden32["col"], int(golden32["P"]), int(golden32["L"]), 0.437)
    assert dec == int(golden32["ird"])
