"""Python face of the CPU oracle (liboracle.so) and of the reference build
(oracle/_ref/libvnecref.so, when present).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The shipped library never
loads anything from this directory.

All structures use the reference's conventions: coords are (x = rows
consumed, y = nonzeros consumed); `win` concatenates each partition's
sorted distinct columns (use_x_indices trimmed to NEC_NUM).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libvnecref.so")

_o = None
_r = None
vp = C.c_void_p
i64 = C.c_int64


def _ptr(a):
    """Pointer to a's data that keeps a alive for the call: callers pass
    temporaries (np.ascontiguousarray of a strided view makes a copy), and
    a bare address would let that copy be freed before the C code reads it."""
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        o = C.CDLL(ORACLE_SO)
        sig = {
            "orc_merge_path": (None, [i64, vp, i64, i64, vp]),
            "orc_items_per_part": (i64, [i64, i64, i64]),
            "orc_partition": (None, [vp, i64, i64, i64, vp, vp]),
            "orc_partition_snap": (None, [vp, i64, i64, i64, vp, vp]),
            "orc_ecr": (None, [vp, i64, vp, vp, vp, vp, vp, vp]),
            "orc_d_build": (i64, [vp, vp, vp, i64, i64, i64, vp, vp, vp, vp, vp]),
            "orc_ird": (C.c_int, [vp, vp, i64, i64, i64, i64, C.c_double, vp]),
            "orc_s_build": (None, [vp, vp, vp, i64, i64, i64, i64, vp, vp, vp, vp, vp, vp,
                                   vp, vp, vp, vp]),
            "orc_spmv_serial": (None, [vp, vp, vp, vp, i64, vp]),
            "orc_spmv_serial_f32": (None, [vp, vp, vp, vp, i64, vp]),
            "orc_spmv_abs": (None, [vp, vp, vp, vp, i64, vp, vp]),
            "orc_spmm_serial": (None, [vp, vp, vp, vp, i64, i64, vp]),
            "orc_spmv_L": (None, [vp, vp, vp, vp, i64, i64, vp, vp, vp, C.c_int]),
            "orc_spmv_D": (None, [vp, vp, vp, i64, i64, i64, vp, vp, vp, vp, vp, vp, vp, vp,
                                  vp, vp, C.c_int]),
            "orc_spmv_S": (None, [vp, i64, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                  C.c_int]),
            "orc_num_threads": (C.c_int, []),
            "orc_feature_ints": (None, [vp, vp, i64, i64, vp]),
        }
        for k, (res, args) in sig.items():
            f = getattr(o, k)
            f.restype = res
            f.argtypes = args
        _o = o
    return _o


def nthreads() -> int:
    return int(lib().orc_num_threads())


# ------------------------------------------------------- selector statistics
def feature_ints(ptr, col, ncols: int) -> np.ndarray:
    """The selector's 32 integer statistics (selector_oracle.c)."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    out = np.zeros(32, dtype=np.uint64)
    lib().orc_feature_ints(_ptr(ptr), _ptr(col), len(ptr) - 1, ncols, _ptr(out))
    return out


# ------------------------------------------------------------------ a1/a2
def merge_path(d: int, ptr: np.ndarray, nnz: int | None = None) -> tuple[int, int]:
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    nrows = len(ptr) - 1
    nnz = int(ptr[-1]) if nnz is None else nnz
    out = np.zeros(2, dtype=np.int32)
    lib().orc_merge_path(d, _ptr(ptr), nrows, nnz, _ptr(out))
    return int(out[0]), int(out[1])


def partition(ptr: np.ndarray, P: int, snap: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """Returns (start, end), each int32 [P, 2] of (x, y)."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    nrows = len(ptr) - 1
    s = np.zeros((P, 2), dtype=np.int32)
    e = np.zeros((P, 2), dtype=np.int32)
    f = lib().orc_partition_snap if snap else lib().orc_partition
    f(_ptr(ptr), nrows, int(ptr[-1]), P, _ptr(s), _ptr(e))
    return s, e


def items_per_part(nrows: int, nnz: int, P: int) -> int:
    return int(lib().orc_items_per_part(nrows, nnz, P))


# ------------------------------------------------------------------ a3
@dataclass
class Ecr:
    ecr: np.ndarray      # int32 [nnz]
    nec: np.ndarray      # int32 [P]
    win_off: np.ndarray  # int64 [P + 1]
    win: np.ndarray      # int32 [sum nec]


def ecr(col: np.ndarray, start: np.ndarray, end: np.ndarray) -> Ecr:
    col = np.ascontiguousarray(col, dtype=np.int32)
    P = len(start)
    nnz = len(col)
    e = np.zeros(max(nnz, 1), dtype=np.int32)
    nec = np.zeros(P, dtype=np.int32)
    woff = np.zeros(P + 1, dtype=np.int64)
    win = np.zeros(max(nnz, 1), dtype=np.int32)
    lib().orc_ecr(_ptr(col), P, _ptr(np.ascontiguousarray(start)), _ptr(np.ascontiguousarray(end)),
                  _ptr(e), _ptr(nec), _ptr(woff), _ptr(win))
    return Ecr(e[:nnz], nec, woff, win[: int(woff[-1])])


# ------------------------------------------------------------------ a4
@dataclass
class DPlan:
    start: np.ndarray
    end: np.ndarray
    ecr: Ecr
    v_row_ptr: np.ndarray  # int32 [nrows + 1]
    col_start: np.ndarray  # int32 [G]
    val_align: np.ndarray  # float64 [G * L]
    L: int

    @property
    def groups(self) -> int:
        return len(self.col_start)


def d_build(ptr, col, val, P: int, L: int) -> DPlan:
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    nrows = len(ptr) - 1
    s, e = partition(ptr, P)
    ec = ecr(col, s, e)
    G = lib().orc_d_build(_ptr(ptr), None, _ptr(ec.ecr), nrows, P, L, _ptr(s), _ptr(e), None,
                          None, None)
    vrp = np.zeros(nrows + 1, dtype=np.int32)
    cs = np.zeros(max(G, 1), dtype=np.int32)
    va = np.zeros(max(G * L, 1), dtype=np.float64)
    val = None if val is None else np.ascontiguousarray(val, dtype=np.float64)
    lib().orc_d_build(_ptr(ptr), _ptr(val), _ptr(ec.ecr), nrows, P, L, _ptr(s), _ptr(e),
                      _ptr(vrp), _ptr(cs), _ptr(va) if val is not None else None)
    return DPlan(s, e, ec, vrp, cs[:G], va[: G * L], L)


def ird(ptr, col, P: int = 36, L: int = 8, thr: float = 0.477) -> tuple[int, float]:
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    r = np.zeros(1, dtype=np.float32)
    dec = lib().orc_ird(_ptr(ptr), _ptr(col), len(ptr) - 1, int(ptr[-1]), P, L, thr, _ptr(r))
    return int(dec), float(r[0])


# ------------------------------------------------------------------ a7/a8
@dataclass
class SPlan:
    start: np.ndarray
    end: np.ndarray
    ecr: Ecr
    tasks: np.ndarray
    spvv_len: np.ndarray
    row_begin: np.ndarray
    row_end: np.ndarray
    s_val: np.ndarray
    s_col: np.ndarray
    task_start: np.ndarray
    task_end: np.ndarray
    L: int


def s_build(ptr, col, val, P: int, L: int, cross_max: int = 32) -> SPlan:
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    nrows = len(ptr) - 1
    nnz = int(ptr[-1])
    s, e = partition(ptr, P, snap=True)
    ec = ecr(col, s, e)
    tasks = np.zeros(max(nrows, 1), dtype=np.int32)
    sl = np.zeros(P, dtype=np.int32)
    rb = np.zeros(max(nrows, 1), dtype=np.int32)
    re = np.zeros(max(nrows, 1), dtype=np.int32)
    sv = np.zeros(max(nnz, 1), dtype=np.float64)
    sc = np.zeros(max(nnz, 1), dtype=np.int32)
    ts = np.zeros(P, dtype=np.int32)
    te = np.zeros(P, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    lib().orc_s_build(_ptr(ptr), _ptr(val), _ptr(ec.ecr), nrows, P, L, cross_max, _ptr(s), _ptr(e),
                      _ptr(tasks), _ptr(sl), _ptr(rb), _ptr(re), _ptr(sv), _ptr(sc), _ptr(ts),
                      _ptr(te))
    return SPlan(s, e, ec, tasks[:nrows], sl, rb[:nrows], re[:nrows], sv[:nnz], sc[:nnz], ts, te, L)


# ------------------------------------------------------------------ SpMV
def spmv_serial(ptr, col, val, x) -> np.ndarray:
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    y = np.zeros(len(ptr) - 1, dtype=np.float64)
    lib().orc_spmv_serial(_ptr(ptr), _ptr(np.ascontiguousarray(col, dtype=np.int32)),
                          _ptr(np.ascontiguousarray(val, dtype=np.float64)),
                          _ptr(np.ascontiguousarray(x, dtype=np.float64)), len(ptr) - 1, _ptr(y))
    return y


def spmv_serial_f32(ptr, col, val, x) -> np.ndarray:
    """test_spmv_vnec.hpp:34-54: FP32 values, x and row sums, in order."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    y = np.zeros(len(ptr) - 1, dtype=np.float32)
    lib().orc_spmv_serial_f32(_ptr(ptr), _ptr(np.ascontiguousarray(col, dtype=np.int32)),
                              _ptr(np.ascontiguousarray(val, dtype=np.float32)),
                              _ptr(np.ascontiguousarray(x, dtype=np.float32)), len(ptr) - 1, _ptr(y))
    return y


def check_rows_f32(y, y_ref, ptr, ax, tol=1e-4) -> tuple[bool, float]:
    """FP32 gate: |y - y_ref| <= max(tol, 2 (len_i + 1) 2^-24) (|A||x|)_i.
    Two FP32 summations of row i in different orders each lie within
    (len_i + 1) 2^-24 (|A||x|)_i of the exact value, so the gate is the
    reference's 1e-4 (test_spmv_vnec.cpp:132) wherever that bound is below
    it, and the bound itself on longer rows; returns (ok, worst ratio to
    (|A||x|)_i)."""
    y = np.asarray(y, dtype=np.float64)
    y_ref = np.asarray(y_ref, dtype=np.float64)
    lens = np.diff(np.asarray(ptr, dtype=np.int64)).astype(np.float64)
    t = np.maximum(tol, 2.0 * (lens + 1.0) * 2.0 ** -24)
    err = np.abs(y - y_ref)
    scale = np.where(ax > 0, ax, 1.0)
    ok = bool(np.all(err <= t * ax) and np.all(np.isfinite(y)))
    return ok, float((err / scale).max() if len(err) else 0.0)


def spmv_abs(ptr, col, val, x) -> tuple[np.ndarray, np.ndarray]:
    """((|A||x|)_i, long-double row sums)."""
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    n = len(ptr) - 1
    ax = np.zeros(n, dtype=np.float64)
    yl = np.zeros(n, dtype=np.float64)
    lib().orc_spmv_abs(_ptr(ptr), _ptr(np.ascontiguousarray(col, dtype=np.int32)),
                       _ptr(np.ascontiguousarray(val, dtype=np.float64)),
                       _ptr(np.ascontiguousarray(x, dtype=np.float64)), n, _ptr(ax), _ptr(yl))
    return ax, yl


def spmm_serial(ptr, col, val, X, k) -> np.ndarray:
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    n = len(ptr) - 1
    Y = np.zeros(n * k, dtype=np.float64)
    lib().orc_spmm_serial(_ptr(ptr), _ptr(np.ascontiguousarray(col, dtype=np.int32)),
                          _ptr(np.ascontiguousarray(val, dtype=np.float64)),
                          _ptr(np.ascontiguousarray(X, dtype=np.float64)), n, k, _ptr(Y))
    return Y


def spmv_L(ptr, col, val, x, P: int, nthreads_: int = 0) -> np.ndarray:
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    n = len(ptr) - 1
    s, e = partition(ptr, P)
    y = np.zeros(n, dtype=np.float64)
    lib().orc_spmv_L(_ptr(ptr), _ptr(np.ascontiguousarray(col, dtype=np.int32)),
                     _ptr(np.ascontiguousarray(val, dtype=np.float64)),
                     _ptr(np.ascontiguousarray(x, dtype=np.float64)), n, P, _ptr(s), _ptr(e),
                     _ptr(y), nthreads_ or nthreads())
    return y


def spmv_D(ptr, val, x, plan: DPlan, nthreads_: int = 0) -> np.ndarray:
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    n = len(ptr) - 1
    y = np.zeros(n, dtype=np.float64)
    e = plan.ecr
    lib().orc_spmv_D(_ptr(ptr), _ptr(np.ascontiguousarray(val, dtype=np.float64)),
                     _ptr(np.ascontiguousarray(x, dtype=np.float64)), n, len(plan.start), plan.L,
                     _ptr(plan.start), _ptr(plan.end), _ptr(e.ecr), _ptr(e.nec), _ptr(e.win_off),
                     _ptr(e.win), _ptr(plan.v_row_ptr), _ptr(plan.col_start), _ptr(plan.val_align),
                     _ptr(y), nthreads_ or nthreads())
    return y


def spmv_S(x, plan: SPlan, nrows: int, nthreads_: int = 0) -> np.ndarray:
    y = np.zeros(nrows, dtype=np.float64)
    e = plan.ecr
    lib().orc_spmv_S(_ptr(np.ascontiguousarray(x, dtype=np.float64)), len(plan.start), plan.L,
                     _ptr(e.nec), _ptr(e.win_off), _ptr(e.win), _ptr(plan.tasks),
                     _ptr(plan.spvv_len), _ptr(plan.row_begin), _ptr(plan.row_end),
                     _ptr(plan.s_val), _ptr(plan.s_col), _ptr(plan.task_start),
                     _ptr(plan.task_end), _ptr(y), nthreads_ or nthreads())
    return y


def check_rows(y, y_ref, ax, tol=1e-12) -> tuple[bool, float]:
    """Per-row gate |y - y_ref| <= tol * (|A||x|)_i; returns (ok, worst ratio)."""
    y = np.asarray(y, dtype=np.float64)
    err = np.abs(y - y_ref)
    scale = np.where(ax > 0, ax, 1.0)
    ratio = err / scale
    ok = bool(np.all(err <= tol * ax + 0.0) and np.all(np.isfinite(y)))
    return ok, float(ratio.max() if len(ratio) else 0.0)


# ------------------------------------------------------------- reference
def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _r
    if _r is None:
        r = C.CDLL(REF_SO)
        sig = {
            "ref_num_threads": (C.c_int, []),
            "ref_lanes_f64": (C.c_int, []),
            "ref_matrix": (vp, [i64, i64, vp, vp, vp]),
            "ref_matrix_free": (None, [vp]),
            "ref_ird": (C.c_int, [vp]),
            "ref_build_d": (vp, [vp]),
            "ref_d_groups": (i64, [vp, vp]),
            "ref_d_export": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "ref_spmv_d": (None, [vp, vp, vp, vp, C.c_int]),
            "ref_free_d": (None, [vp]),
            "ref_build_s": (vp, [vp]),
            "ref_s_export": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "ref_spmv_s": (None, [vp, vp, vp, vp, C.c_int]),
            "ref_free_s": (None, [vp]),
            "ref_build_l": (vp, [vp]),
            "ref_l_export": (None, [vp, vp]),
            "ref_spmv_l": (None, [vp, vp, vp, vp, C.c_int]),
            "ref_free_l": (None, [vp]),
            "ref_time_spmv": (C.c_double, [C.c_int, vp, vp, vp, vp, C.c_int]),
            "ref_lanes_f32": (C.c_int, []),
            "ref_matrix32": (vp, [i64, i64, vp, vp, vp]),
            "ref_ird32": (C.c_int, [vp]),
            "ref_build_d32": (vp, [vp]),
            "ref_d32_groups": (i64, [vp, vp]),
            "ref_d32_export": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "ref_spmv_d32": (None, [vp, vp, vp, vp, C.c_int]),
            "ref_free_d32": (None, [vp]),
            "ref_build_s32": (vp, [vp]),
            "ref_s32_export": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "ref_spmv_s32": (None, [vp, vp, vp, vp, C.c_int]),
            "ref_free_s32": (None, [vp]),
            "ref_spmv_l32": (None, [vp, vp, vp, vp, C.c_int]),
        }
        for k, (res, args) in sig.items():
            f = getattr(r, k)
            f.restype = res
            f.argtypes = args
        _r = r
    return _r


class RefMatrix:
    """A reference SpB_Matrix built from a CSR (kept alive with its arrays);
    fp32=True makes an SpB_FP32 matrix and routes every call to the
    reference's FP32 entry points."""

    def __init__(self, nrows, ncols, ptr, col, val, fp32: bool = False):
        self.nrows, self.ncols = nrows, ncols
        self.fp32 = fp32
        self.real = np.float32 if fp32 else np.float64
        self.ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=self.real)
        self.nnz = int(self.ptr[-1])
        mk = ref().ref_matrix32 if fp32 else ref().ref_matrix
        self.h = mk(nrows, ncols, _ptr(self.ptr), _ptr(self.col), _ptr(self.val))

    def close(self):
        if self.h:
            ref().ref_matrix_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ird(self) -> int:
        return int((ref().ref_ird32 if self.fp32 else ref().ref_ird)(self.h))

    def build_d(self) -> dict:
        r = ref()
        P = r.ref_num_threads()
        L = r.ref_lanes_f32() if self.fp32 else r.ref_lanes_f64()
        p = (r.ref_build_d32 if self.fp32 else r.ref_build_d)(self.h)
        G = (r.ref_d32_groups if self.fp32 else r.ref_d_groups)(p, self.h)
        out = dict(coords=np.zeros((P, 4), np.int32), ecr=np.zeros(max(self.nnz, 1), np.int32),
                   nec=np.zeros(P, np.int32), win=np.zeros(max(self.nnz, 1), np.int32),
                   v_row_ptr=np.zeros(self.nrows + 1, np.int32),
                   col_start=np.zeros(max(G, 1), np.int32),
                   val_align=np.zeros(max(G * L, 1), self.real))
        (r.ref_d32_export if self.fp32 else r.ref_d_export)(p, self.h, *[_ptr(out[k]) for k in
                                    ("coords", "ecr", "nec", "win", "v_row_ptr", "col_start",
                                     "val_align")])
        out["ecr"] = out["ecr"][: self.nnz]
        out["win"] = out["win"][: int(out["nec"].sum())]
        out["col_start"] = out["col_start"][:G]
        out["val_align"] = out["val_align"][: G * L]
        out["_plan"] = p
        return out

    def build_s(self) -> dict:
        r = ref()
        P = r.ref_num_threads()
        p = (r.ref_build_s32 if self.fp32 else r.ref_build_s)(self.h)
        n, z = self.nrows, self.nnz
        out = dict(nec=np.zeros(P, np.int32), win=np.zeros(max(z, 1), np.int32),
                   spvv_len=np.zeros(P, np.int32), tasks=np.zeros(max(n, 1), np.int32),
                   row_begin=np.zeros(max(n, 1), np.int32), row_end=np.zeros(max(n, 1), np.int32),
                   s_col=np.zeros(max(z, 1), np.int32), s_val=np.zeros(max(z, 1), self.real),
                   task_se=np.zeros((P, 2), np.int32))
        (r.ref_s32_export if self.fp32 else r.ref_s_export)(p, self.h, *[_ptr(out[k]) for k in
                                    ("nec", "win", "spvv_len", "tasks", "row_begin", "row_end",
                                     "s_col", "s_val", "task_se")])
        out["win"] = out["win"][: int(out["nec"].sum())]
        for k in ("tasks", "row_begin", "row_end"):
            out[k] = out[k][:n]
        out["s_col"] = out["s_col"][:z]
        out["s_val"] = out["s_val"][:z]
        out["_plan"] = p
        return out

    def build_l(self) -> dict:
        r = ref()
        P = r.ref_num_threads()
        p = r.ref_build_l(self.h)
        coords = np.zeros((P, 4), np.int32)
        r.ref_l_export(p, _ptr(coords))
        return dict(coords=coords, _plan=p)

    def spmv(self, variant: str, plan: dict, x, reps: int = 1) -> np.ndarray:
        y = np.zeros(self.nrows, self.real)
        x = np.ascontiguousarray(x, dtype=self.real)
        r = ref()
        if self.fp32:
            f = {"D": r.ref_spmv_d32, "S": r.ref_spmv_s32, "L": r.ref_spmv_l32}[variant]
        else:
            f = {"D": r.ref_spmv_d, "S": r.ref_spmv_s, "L": r.ref_spmv_l}[variant]
        f(plan["_plan"], self.h, _ptr(x), _ptr(y), reps)
        return y

    def time_spmv(self, variant: str, plan: dict, x, reps: int = 40) -> tuple[float, np.ndarray]:
        y = np.zeros(self.nrows, np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = {"D": 0, "S": 1, "L": 2}[variant]
        t = ref().ref_time_spmv(v, plan["_plan"], self.h, _ptr(x), _ptr(y), reps)
        return float(t), y

    def free(self, variant: str, plan: dict):
        r = ref()
        if self.fp32:
            f = {"D": r.ref_free_d32, "S": r.ref_free_s32, "L": r.ref_free_l}[variant]
        else:
            f = {"D": r.ref_free_d, "S": r.ref_free_s, "L": r.ref_free_l}[variant]
        f(plan["_plan"])
