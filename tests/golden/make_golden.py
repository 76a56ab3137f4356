"""Writes the golden fixtures under tests/golden/ from the reference VNEC
library compiled from its own sources (oracle/_ref/libvnecref.so, built by
`make -C oracle ref` in a container where the reference is mounted).

Run from the repository root:  python tests/golden/make_golden.py
(--f32: only the FP32 fixtures, see write_f32)

Each fixture `<name>.npz` holds the input CSR (ptr, col, val, x) and what
the reference computes on it at its compiled parameters (P = NUM_THREADS =
36, L = nLanes_f64 = 8):
  coords (P x 4), ecr, nec, win (concatenated use_x_indices), v_row_ptr,
  col_start, val_align        — Build_VNEC_D_FP64
  s_* (when safe, see below)  — Build_VNEC_S_FP64
  l_coords                    — Build_VNEC_L
  ird                         — IRD_VNEC decision (0 = D, 1 = S)
  y_D / y_S / y_L             — SpMV_VNEC_*_FP64 outputs where the
                                reference's own code is sound on the input
  y_kat                       — closed-form known answer when values follow
                                the reference harness (A_ij = i, x_j = j)

S is exported only when every snapped partition has rows of at least two
lengths and M >= P (P - 1): otherwise Build_VNEC_S_FP64 writes one int past
a bucket or reads ptr[nrows + 1] (spmv_vnec_S_for_mxv.c:457-468, 363-370).
y_D is stored only when no partition lies inside a single row (the D carry
re-reads [ptr[end.x], start.y) otherwise, spmv_vnec_D_for_mxv.c:471), and
y_D / y_L only when no two partitions carry into one row (their fix-up
loop races then, :480-484 and spmv_vnec_L_for_mxv.c:160-164).
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from vnec_amd import matgen as g  # noqa: E402

REF_DATA = "/root/reference/data/graph"
OUT = os.path.dirname(os.path.abspath(__file__))


def kat_values(ptr, ncols):
    """A_ij = i (row index), x_j = j — test_spmv_vnec.cpp:159-186."""
    n = len(ptr) - 1
    val = np.repeat(np.arange(n, dtype=np.float64), np.diff(ptr))
    x = np.arange(ncols, dtype=np.float64)
    return val, x


def s_safe(ptr, P):
    s, e = O.partition(ptr, P, snap=True)
    lens = np.diff(ptr)
    if len(ptr) - 1 + ptr[-1] < P * (P - 1):
        return False
    for t in range(P):
        a, b = int(s[t, 0]), int(e[t, 0])
        if b <= a:
            return False
        seg = lens[a:b]
        if np.all(seg == seg[0]):
            return False
    return True


def d_carry_sound(coords):
    inside = np.any(coords[:, 0] == coords[:, 2])
    ends = coords[:-1, 2]
    shared = len(np.unique(ends)) != len(ends)
    return (not inside) and (not shared)


def write(name, nrows, ncols, ptr, col, val, x, kat=False):
    R = O.RefMatrix(nrows, ncols, ptr, col, val)
    P = O.ref().ref_num_threads()
    d = R.build_d()
    out = dict(nrows=np.int64(nrows), ncols=np.int64(ncols), ptr=ptr.astype(np.int64),
               col=col.astype(np.int32), val=val.astype(np.float64), x=x.astype(np.float64),
               P=np.int64(P), L=np.int64(O.ref().ref_lanes_f64()))
    for k in ("coords", "ecr", "nec", "win", "v_row_ptr", "col_start", "val_align"):
        out[k] = d[k]
    out["ird"] = np.int64(R.ird())
    lp = R.build_l()
    out["l_coords"] = lp["coords"]
    shared = len(np.unique(d["coords"][:-1, 2])) != P - 1
    if d_carry_sound(d["coords"]):
        out["y_D"] = R.spmv("D", d, x)
    if not shared:
        out["y_L"] = R.spmv("L", lp, x)
    R.free("D", d)
    R.free("L", lp)
    if s_safe(ptr, P):
        s = R.build_s()
        for k in ("nec", "win", "spvv_len", "tasks", "row_begin", "row_end", "s_col", "s_val",
                  "task_se"):
            out["s_" + k] = s[k]
        out["y_S"] = R.spmv("S", s, x)
        R.free("S", s)
    if kat:
        rows = np.repeat(np.arange(nrows, dtype=np.float64), np.diff(ptr))
        y = np.zeros(nrows)
        np.add.at(y, np.repeat(np.arange(nrows), np.diff(ptr)), rows * col.astype(np.float64))
        out["y_kat"] = y
    R.close()
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)
    print(name, nrows, int(ptr[-1]), "S" if "s_tasks" in out else "-",
          "yD" if "y_D" in out else "-", "yL" if "y_L" in out else "-")


def write_f32(name, nrows, ncols, ptr, col, val, x, kat=False):
    """`<name>_f32.npz`: the same input with FP32 values and x, and what the
    reference's FP32 entry points compute on it (L = nLanes_f32 = 16):
    Build_VNEC_D_FP32 / Build_VNEC_S_FP32 layouts, IRD_VNEC(A, SpB_FP32),
    SpMV_VNEC_*_FP32 outputs (under the same soundness rules as FP64, and
    only where finite: the FP32 D kernel also reads past NEC_NUM,
    spmv_vnec_D_for_mxv.c:206)."""
    val32 = np.asarray(val, np.float32)
    x32 = np.asarray(x, np.float32)
    R = O.RefMatrix(nrows, ncols, ptr, col, val32, fp32=True)
    P = O.ref().ref_num_threads()
    d = R.build_d()
    out = dict(nrows=np.int64(nrows), ncols=np.int64(ncols), ptr=ptr.astype(np.int64),
               col=col.astype(np.int32), val=val32, x=x32, P=np.int64(P),
               L=np.int64(O.ref().ref_lanes_f32()))
    for k in ("coords", "ecr", "nec", "win", "v_row_ptr", "col_start", "val_align"):
        out[k] = d[k]
    out["ird"] = np.int64(R.ird())
    lp = R.build_l()
    shared = len(np.unique(d["coords"][:-1, 2])) != P - 1
    if d_carry_sound(d["coords"]):
        y = R.spmv("D", d, x32)
        if np.all(np.isfinite(y)):
            out["y_D"] = y
    if not shared:
        out["y_L"] = R.spmv("L", lp, x32)
    R.free("D", d)
    R.free("L", lp)
    if s_safe(ptr, P):
        s = R.build_s()
        for k in ("nec", "win", "spvv_len", "tasks", "row_begin", "row_end", "s_col", "s_val",
                  "task_se"):
            out["s_" + k] = s[k]
        out["y_S"] = R.spmv("S", s, x32)
        R.free("S", s)
    if kat:
        rows = np.repeat(np.arange(nrows, dtype=np.float64), np.diff(ptr))
        y = np.zeros(nrows)
        np.add.at(y, np.repeat(np.arange(nrows), np.diff(ptr)), rows * col.astype(np.float64))
        out["y_kat"] = y
    R.close()
    np.savez_compressed(os.path.join(OUT, name + "_f32.npz"), **out)
    print(name + "_f32", nrows, int(ptr[-1]), "S" if "s_tasks" in out else "-",
          "yD" if "y_D" in out else "-", "yL" if "y_L" in out else "-")


def main_f32():
    for fx in ("pdf_test", "grb_ppt"):
        nr, nc, ptr, col, _ = g.read_mtx(os.path.join(REF_DATA, fx, fx + ".mtx"))
        val, x = kat_values(ptr, nc)
        write_f32(fx + "_kat", nr, nc, ptr, col, val, x, kat=True)
    m = g.rmat(11, 16, seed=7)
    val, x = kat_values(m.ptr, m.ncols)
    write_f32("rmat11_kat", m.nrows, m.ncols, m.ptr, m.col, val, x, kat=True)
    for name, m in [("stencil10", g.stencil27(10, seed=2)),
                    ("uniform2000_k9", g.uniform(2000, 9, seed=8)),
                    ("powerlaw2500_a2", g.powerlaw(2500, 2.0, 2, 800, seed=6)),
                    ("banded3000_hb5", g.banded(3000, 5, seed=4))]:
        write_f32(name, m.nrows, m.ncols, m.ptr, m.col, m.val, m.x)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libvnecref.so missing: run `make -C oracle ref`")
    for fx in ("pdf_test", "grb_ppt"):
        nr, nc, ptr, col, _ = g.read_mtx(os.path.join(REF_DATA, fx, fx + ".mtx"))
        val, x = kat_values(ptr, nc)
        write(fx + "_kat", nr, nc, ptr, col, val, x, kat=True)
    m = g.rmat(11, 16, seed=7)
    val, x = kat_values(m.ptr, m.ncols)
    write("rmat11_kat", m.nrows, m.ncols, m.ptr, m.col, val, x, kat=True)
    cases = [
        ("rmat10", g.rmat(10, 16, seed=3)),
        ("stencil10", g.stencil27(10, seed=2)),
        ("banded3000_hb5", g.banded(3000, 5, seed=4)),
        ("blockdiag2000_bs16", g.blockdiag(2000, 16, seed=5)),
        ("powerlaw2500_a2", g.powerlaw(2500, 2.0, 2, 800, seed=6)),
        ("uniform2000_k9", g.uniform(2000, 9, seed=8)),
    ]
    for name, m in cases:
        write(name, m.nrows, m.ncols, m.ptr, m.col, m.val, m.x)


if __name__ == "__main__":
    if sys.argv[1:] == ["--f32"]:
        main_f32()
    else:
        main()
        main_f32()
