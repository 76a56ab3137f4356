#!/usr/bin/env python3
"""bench.py — FP64 SpMV y = A x on one MI355X, on the 27-point stencil of a
128^3 grid (2,097,152 rows, 55,742,968 nonzeros, CSR int32, FP64 values,
seed 1): the configuration the headline metric is quoted on.

A "step" is one SpMV over the whole matrix with A, x and y resident in HBM.
`value` is algorithmic GB/s over the timed steps, with
  B = (vb + 4) nnz + 4 (nrows + 1) + 8 ncols + 8 nrows   (vb = 8 for FP64)
bytes per SpMV, whatever the on-device format moves.  The roofline object
uses the same B over the SpMV kernel's mean duration (HIP events on the
library's stream).  The CPU baseline is the reference VNEC code (compiled
from its sources into oracle/_ref) with its own dispatch, timed like the
reference harness (1 warm-up + 40 SpMVs), on this host's cores.

  python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNEL_NAME = {"S": "k_spmv_slice", "L": "k_spmv_stream", "E": "k_spmv_estream"}
METRIC = "FP64 SpMV GB/s as % of HBM peak; learned-vs-plain speedup at equal error"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class _QuietStdout:
    """Silences fd 1 (the reference library printf()s its IRD ratio), so
    the bench's JSON line is the only line on stdout."""

    def __enter__(self):
        import ctypes
        self.libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)
        return self

    def __exit__(self, *exc):
        self.libc.fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)
        return False


def cpu_baseline(m, y_gpu):
    """Reference VNEC CPU path on this host (the checker's leg)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # noqa: E402  (checker / baseline only)

    B = m.algorithmic_bytes(8)
    flags = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    out = {}
    if O.ref_available() and "avx512f" in flags:
        R = O.RefMatrix(m.nrows, m.ncols, m.ptr, m.col, m.val)
        threads = O.ref().ref_num_threads()
        t0 = time.perf_counter()
        if m.nrows < 3_000_000:  # dispatch of test_spmv_vnec.cpp:197-256
            kind = "D" if R.ird() == 0 else "S"
        else:
            kind = "L"
        plan = {"D": R.build_d, "S": R.build_s, "L": R.build_l}[kind]()
        build_s = time.perf_counter() - t0
        reps = 40
        t_spmv, y_cpu = R.time_spmv(kind, plan, m.x, reps)
        R.free(kind, plan)
        R.close()
        out = {"value": B / t_spmv / 1e9, "unit": "GB/s", "cores": int(threads),
               "kind": "reference",
               "sample": (f"full workload, reference VNEC-{kind} (IRD dispatch) compiled "
                          f"-O3 -march=skylake-avx512, OpenMP {threads} threads "
                          f"(NUM_THREADS) on {os.cpu_count()} visible host CPUs; "
                          f"1 warm-up + {reps} SpMVs; build {build_s:.1f} s untimed"),
               "ms_per_spmv": t_spmv * 1e3}
    else:
        nt = O.nthreads()
        P = max(nt, 1)
        O.spmv_L(m.ptr, m.col, m.val, m.x, P, nt)
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            y_cpu = O.spmv_L(m.ptr, m.col, m.val, m.x, P, nt)
        t_spmv = (time.perf_counter() - t0) / reps
        out = {"value": B / t_spmv / 1e9, "unit": "GB/s", "cores": int(nt), "kind": "port",
               "sample": f"full workload, oracle merge-path CSR (VNEC-L restatement), "
                         f"{nt} OpenMP threads, {reps} SpMVs (reference build unavailable)",
               "ms_per_spmv": t_spmv * 1e3}
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    err = np.abs(y_gpu - y_cpu) / np.where(ax > 0, ax, 1.0)
    out["gpu_vs_cpu_max_rel_row_err"] = float(err.max())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the SpMV would shard by merge-path diagonal ranges; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    from vnec_amd import matgen as g
    from vnec_amd import spb

    if spb.device_count() < 1:
        raise SystemExit("no GPU visible")
    log(f"generating 27-point stencil {args.grid}^3 ...")
    m = g.stencil27(args.grid, seed=1)
    B = m.algorithmic_bytes(8)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    dx = spb.DeviceArray.from_array(m.x)
    dy = spb.DeviceArray(8 * m.nrows)

    log("building plan (selector) ...")
    t0 = time.perf_counter()
    plan = spb.Plan(A, variant=None)
    build_ms = (time.perf_counter() - t0) * 1e3
    info = plan.info()
    log(f"plan: {info}")

    for _ in range(args.warmup):
        plan.spmv_dev(dx.ptr, dy.ptr)
    spb.L.load().vnec_sync_ext()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.spmv_dev(dx.ptr, dy.ptr)
    spb.L.check(spb.L.load().vnec_sync_ext(), "sync")
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    value = B / (ms_step / 1e3) / 1e9
    y_gpu = dy.to_array(np.float64, m.nrows)

    k_ms, fix_ms = plan.time_kernels(dx.ptr, dy.ptr, max(args.steps // 2, 20))
    graph_ms = plan.time(dx.ptr, dy.ptr, max(args.steps, 20), graph=True)
    achieved = B / (k_ms / 1e3) / 1e9
    log(f"wall {ms_step * 1e3:.1f} us/step  kernel {k_ms * 1e3:.1f} us  fixup "
        f"{fix_ms * 1e3:.1f} us  graph {graph_ms * 1e3:.1f} us")

    variants = {}
    for name, var, ipt, enc, vb in (("S_fp64", "S", 3, "fp64", 8), ("S_hilo", "S", 3, "hilo", 8),
                                    ("S_hi", "S", 3, "hi", 4), ("L_fp64", "L", 7, "fp64", 8),
                                    ("E_fp64", "E", 5, "fp64", 8)):
        p = spb.Plan(A, var, ipt, encoding=enc)
        km, _ = p.time_kernels(dx.ptr, dy.ptr, 50)
        gm = p.time(dx.ptr, dy.ptr, 100, graph=True)
        bv = m.algorithmic_bytes(vb)
        variants[name] = {"kernel_us": round(km * 1e3, 2), "spmv_us": round(gm * 1e3, 2),
                          "GBps": round(bv / (gm / 1e3) / 1e9, 1), "bytes": bv}
        p.free()
    plan.spmv_dev(dx.ptr, dy.ptr)
    spb.L.load().vnec_sync_ext()

    # measured streaming-copy ceiling (also the byte-counter calibration kernel)
    import ctypes as C
    cp = C.c_double()
    spb.L.check(spb.L.load().vnec_copy_bandwidth_ext(1 << 30, 10, C.byref(cp)), "copy bandwidth")

    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_stencil27_128.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        with _QuietStdout():
            cpu = cpu_baseline(m, y_gpu)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"27-point stencil {args.grid}^3 FP64 SpMV (CSR int32, seed 1)",
                   "nrows": m.nrows, "nnz": m.nnz, "algorithmic_bytes": B,
                   "variant": info.variant, "ipt": info.ipt, "tiles": info.ntiles,
                   "parallelism": "single GPU"},
        "pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
        "gflops": round(2.0 * m.nnz / (ms_step / 1e3) / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": KERNEL_NAME[info.variant],
                     "kernel_us": round(k_ms * 1e3, 2), "fixup_us": round(fix_ms * 1e3, 2)},
        "graph_spmv_us": round(graph_ms * 1e3, 2),
        "copy_peak_gbs": round(cp.value, 1),
        "pct_of_copy_peak": round(100.0 * achieved / cp.value, 2),
        "plan_build_ms": round(build_ms, 1),
        "variants": variants,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
