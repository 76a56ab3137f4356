#!/usr/bin/env python3
"""bench.py — FP64 SpMV y = A x on one MI355X.

Default run: the 27-point stencil of a 128^3 grid (2,097,152 rows,
55,742,968 nonzeros, CSR int32, FP64 values, seed 1) as the headline line,
and R-MAT scale 18 (262,144 rows, edge factor 16, seed 1) in
`workloads.rmat18` -- the two configurations the HBM target is quoted on.
Other workloads of the suite are selectable with --workload (rmat18, skew,
batch, batch8).

A "step" is one SpMV over the whole workload with A, x and y resident in
HBM.  Two byte counts are reported for every plan:
  format bytes -- what the plan's own layout must move per SpMV (every
      array its kernels read counted once, x read once, y written once;
      vnec_plan_bytes_ext).  `value`, `pct_hbm_peak` and `roofline.frac`
      use these, so they are fractions of the HBM peak (<= 1).
  CSR bytes -- the format-agnostic figure, CSR with int32 columns:
      B = (vb + 4) nnz + 4 (nrows + 1) + 8 k ncols + 8 k nrows (vb = 8 for
      FP64), reported beside as `csr_equivalent_GBps`.  A format that reads
      no column indices (pattern slices) moves fewer bytes than B, so its
      CSR-equivalent rate can pass the HBM peak.
The roofline object divides the main kernel's format bytes by its mean
duration (back-to-back launches between one pair of HIP events on the
library's stream).  `traffic` is the counter figure (FETCH_SIZE and
WRITE_SIZE, rocprofv3) stored under profiles/ for the same plan shape: bytes
between L2 and the memory fabric, an upper bound on HBM bytes because
Infinity Cache hits count too.  The CPU baseline is the reference VNEC code
(compiled from its sources into oracle/_ref) with its own dispatch, timed
like the reference harness (1 warm-up + 40 SpMVs), on this host's CPUs;
beside it, the oracle's merge-path restatement with one partition per host
CPU.

  python bench.py [--workload stencil] [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# dominant kernel of each plan shape: (variant, persistent schedule?)
KERNEL_NAME = {("S", False): "k_spmv_slice", ("L", True): "k_spmv_stream",
               ("E", True): "k_spmv_estream", ("L", False): "k_spmv_merge1",
               ("E", False): "k_spmv_merge1", ("G", False): "k_spmv_pattern",
               ("D", False): "k_spmv_d", ("H", True): "k_spmv_hot4", ("C", True): "k_spmv_stream",
               ("C", False): "k_spmv_merge1"}
PLAIN_CANDIDATE = 4  # L, 7 items per thread, persistent: the plain merge-path CSR kernel
METRIC = "FP64 SpMV GB/s as % of HBM peak; learned-vs-plain speedup at equal error"
R18_FLUSH = 1 << 29  # 512 MiB read between SpMVs (> L2 + 256 MB Infinity Cache): an HBM figure


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class _QuietStdout:
    """Silences fd 1 (the reference library printf()s its IRD ratio), so
    the bench's JSON line is the only line on stdout."""

    def __enter__(self):
        self.libc = C.CDLL(None)
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


def workload(name, g):
    """(matrix, description, right-hand sides)."""
    if name == "stencil":
        m = g.stencil27(128, seed=1)
        return m, "27-point stencil 128^3 FP64 SpMV (CSR int32, seed 1)", 1
    if name == "rmat18":
        m = g.rmat(18, 16, 0.57, 0.19, 0.19, seed=1)
        return m, "R-MAT scale 18, edge factor 16, (.57,.19,.19,.05), seed 1", 1
    if name == "skew":
        m = g.skew(22, 16, 8, 500_000, seed=1)
        return m, "R-MAT scale 22 ef 16 + 8 dense rows x 500k distinct columns, seed 1", 1
    if name in ("batch", "batch8"):
        k = 8 if name == "batch8" else 1
        m = g.batch(512, 1000, 200_000, 4.0, 64.0, nrhs=k, seed=1)
        return m, (f"512 independent matrices, rows log-uniform [1k, 200k], nnz/row "
                   f"Poisson(mu ~ U[4, 64]), per-matrix x, k = {k}, one launch"), k
    raise SystemExit(f"unknown workload {name}")


def host_cpus():
    """CPUs this process may run on (the box's share, whatever the machine has)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(m, k, y_gpu, variant_y):
    """Reference VNEC CPU path on this host (the checker's leg), and the
    oracle's merge-path restatement with one partition per host CPU.  Also
    reports, for each extra GPU result in variant_y (one right-hand side),
    max_i |y_i - y_ref,i| / (|A||x|)_i against the reference's FP64 y."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # noqa: E402  (checker / baseline only)

    B = m.algorithmic_bytes(8, k)
    cpus = host_cpus()
    flags = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    col = m.col if m.col_off is None else (
        m.col + np.repeat(m.col_off[:-1], np.diff(m.ptr[m.row_off])).astype(np.int32))
    xs = [np.ascontiguousarray(m.x[c::k]) for c in range(k)]

    # the oracle's VNEC-L restatement at P = host CPUs, one OpenMP thread each
    nt = O.nthreads()
    reps = 10
    t_port = 0.0
    ys_port = []
    for c in range(k):
        O.spmv_L(m.ptr, col, m.val, xs[c], nt, nt)
        t0 = time.perf_counter()
        for _ in range(reps):
            y = O.spmv_L(m.ptr, col, m.val, xs[c], nt, nt)
        t_port += (time.perf_counter() - t0) / reps
        ys_port.append(y)
    port = {"value": round(B / t_port / 1e9, 2), "unit": "GB/s", "cores": int(nt), "kind": "port",
            "sample": (f"full workload, the oracle's merge-path CSR (VNEC-L restatement, "
                       f"-O3 -march=x86-64-v3) with P = {nt} partitions on {nt} OpenMP threads, "
                       f"{cpus} host CPUs visible; 1 warm-up + {reps} SpMVs"
                       + (f" per right-hand side (k = {k})" if k > 1 else "")),
            "ms_per_spmv": round(t_port * 1e3, 4)}

    ys = ys_port
    out = port
    if O.ref_available() and "avx512f" in flags:
        R = O.RefMatrix(m.nrows, m.ncols, m.ptr, col, m.val)
        threads = O.ref().ref_num_threads()
        t0 = time.perf_counter()
        if m.nrows < 3_000_000:  # dispatch of test_spmv_vnec.cpp:197-256
            kind = "D" if R.ird() == 0 else "S"
        else:
            kind = "L"
        plan = {"D": R.build_d, "S": R.build_s, "L": R.build_l}[kind]()
        build_s = time.perf_counter() - t0
        reps = 40
        t_spmv = 0.0
        ys = []
        for c in range(k):
            t, y = R.time_spmv(kind, plan, xs[c], reps)
            t_spmv += t
            ys.append(y)
        R.free(kind, plan)
        R.close()
        out = {"value": round(B / t_spmv / 1e9, 2), "unit": "GB/s", "cores": cpus,
               "threads": int(threads), "kind": "reference",
               "sample": (f"full workload, reference VNEC-{kind} (its dispatch) compiled "
                          f"-O3 -march=skylake-avx512; its OpenMP team is fixed at compile time "
                          f"(NUM_THREADS = {threads}, SpB_mxv.h:28) and ran on the {cpus} host "
                          f"CPUs visible; 1 warm-up + {reps} SpMVs"
                          + (f" per right-hand side (k = {k}, no multi-RHS in the reference)"
                             if k > 1 else "") + f"; build {build_s:.1f} s untimed"),
               "ms_per_spmv": round(t_spmv * 1e3, 4),
               "port_at_host_cpus": port}
    worst = 0.0
    for c in range(k):
        ax, _ = O.spmv_abs(m.ptr, col, m.val, xs[c])
        err = np.abs(y_gpu[c::k] - ys[c]) / np.where(ax > 0, ax, 1.0)
        worst = max(worst, float(err.max()))
    out["gpu_vs_cpu_max_rel_row_err"] = worst
    ax, _ = O.spmv_abs(m.ptr, col, m.val, xs[0])
    scale = np.where(ax > 0, ax, 1.0)
    out["variant_errors"] = {name: float((np.abs(y - ys[0]) / scale).max())
                             for name, y in variant_y.items()}
    return out


def _abs_rows(ptr, col, val, x):
    rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    return np.bincount(rows, weights=np.abs(val) * np.abs(x[col]), minlength=len(ptr) - 1)


def selector_eval(spb, n):
    """Learned selector vs the plain kernel, the reference dispatch and the
    exhaustive autotuner, live, on n held-out selector matrices."""
    from concurrent.futures import ThreadPoolExecutor

    from vnec_amd.selector import specs

    idx = [specs.N_TRAIN + (8 * j) % specs.N_HELD for j in range(n)]
    pool = ThreadPoolExecutor(max_workers=8)
    futs = [pool.submit(lambda i: specs.generate(specs.spec(i)), i) for i in idx]
    rows = []
    worst = 0.0
    for i, f in zip(idx, futs):
        m = f.result()
        A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
        t, best = spb.autotune(A, 20)
        learned = spb.select(A, "learned")
        ref = spb.select(A, "reference")
        rule = spb.select(A, "rule")
        ys = []
        for c in (learned, PLAIN_CANDIDATE):
            ci = spb.candidates()[c]
            p = spb.Plan(A, ci.variant, ci.ipt, grid=ci.grid)
            xv, yv = spb.Vector.from_array(m.x), spb.Vector(m.nrows)
            p.spmv(yv, xv)
            ys.append(yv.to_array())
            p.free()
        ax = _abs_rows(m.ptr, m.col, m.val, m.x)
        worst = max(worst, float((np.abs(ys[0] - ys[1]) / np.where(ax > 0, ax, 1.0)).max()))
        rows.append((t[best], t[learned], t[PLAIN_CANDIDATE], t[ref], t[rule]))
        log(f"selector [{i}] {specs.spec(i)['family']:9s} nnz {m.nnz:>9d} best {best:2d} "
            f"learned {learned:2d} ({t[learned] / t[best]:.3f}x of best)")
        A.free()
    pool.shutdown()
    r = np.array(rows)
    geo = lambda a, b: float(np.exp(np.mean(np.log(a / b))))  # noqa: E731
    return {"matrices": n,
            "sample": (f"held-out selector matrices {idx[0]}..{max(idx)} (every 8th of 200: "
                       "R-MAT, banded, block-diagonal, uniform, power-law), timed live with "
                       "hipGraph replay; plain = merge-path CSR kernel, 7 items/thread, persistent"),
            "learned_vs_plain_speedup_geomean": round(geo(r[:, 2], r[:, 1]), 4),
            "learned_vs_plain_speedup_aggregate": round(float(r[:, 2].sum() / r[:, 1].sum()), 4),
            "learned_vs_reference_dispatch_speedup_geomean": round(geo(r[:, 3], r[:, 1]), 4),
            "learned_vs_rule_speedup_geomean": round(geo(r[:, 4], r[:, 1]), 4),
            "learned_regret_vs_exhaustive_geomean": round(geo(r[:, 1], r[:, 0]), 4),
            "learned_regret_vs_exhaustive_aggregate": round(float(r[:, 1].sum() / r[:, 0].sum()), 4),
            "max_rel_row_diff_learned_vs_plain": worst}


def selection_label(spb, A, mode, info, k):
    """How the plan was chosen: forced, the MLP's own pick, or the MLP's pick
    refined by timing its neighbours (beyond the model's training sizes)."""
    if mode is None:
        return "forced"
    if mode == "rule" and info.variant == "G":
        return "rule (pattern rule)"
    if mode != "learned" or k > 1:
        return mode
    c = spb.candidates()[spb.select(A, "learned")]
    # persistent picks report their workgroup count; per-tile-carry picks (-1)
    # run the strided launch (-2) by themselves once the tiles fill the GPU
    same_grid = c.grid == info.grid or (c.grid >= 0 and info.grid >= 0) or (c.grid < 0 and info.grid < 0)
    flags = info.lanes if info.variant in ("L", "C") and info.lanes != 64 else 0
    extra = ((" 512-thread tiles" if flags & 1024 else "") + (" gathers issued together" if flags & 2048 else "")
             + (" relabelled columns" if flags & 1 else ""))
    if (c.variant, c.ipt, c.lanes) == (info.variant, info.ipt, info.lanes) and same_grid:
        return f"learned (MLP pick: candidate {c.variant}{c.ipt}{extra})"
    form = f"{info.variant}{info.ipt} grid {info.grid}{extra}"
    return f"learned (MLP pick {c.variant}{c.ipt}, measured neighbour kept: {form})"


def kernel_name(info, k):
    if k > 1:
        return "k_spmv_mergeKg" if k >= 4 else "k_spmv_mergeK"
    if info.grid == -2:
        return "k_spmv_strided"
    return KERNEL_NAME.get((info.variant, info.grid >= 0), "?")


def stored_traffic(name, info):
    """Counter bytes per launch stored under profiles/ for this plan shape
    (None when the stored run was of another shape)."""
    tf = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(tf):
        return None
    with open(tf) as f:
        tj = json.load(f)
    if (tj.get("variant"), tj.get("ipt"), tj.get("grid"), tj.get("lanes", info.lanes)) == \
            (info.variant, info.ipt, info.grid, info.lanes):
        return tj.get("hbm_bytes_per_launch")
    return None


def measure(spb, lib, name, args, variant=None, mode=None, steps=None, warmup=None,
            flush=False):
    """Builds the workload's plan and times it.  Returns (record, matrix,
    k, y_gpu, A, dx, dy, plan)."""
    from vnec_amd import matgen as g

    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    log(f"generating {name} ...")
    m, desc, k = workload(name, g)
    B = m.algorithmic_bytes(8, k)
    if m.row_off is not None:
        A = spb.Matrix.from_batch(m.row_off, m.col_off, m.ptr, m.col, m.val)
    else:
        A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    dx = spb.DeviceArray.from_array(m.x)
    dy = spb.DeviceArray(8 * m.nrows * k)

    log(f"building plan ({name}) ...")
    t0 = time.perf_counter()
    if variant:
        plan = spb.Plan(A, variant, args.ipt or (3 if variant == "S" else 7), nrhs=k, grid=args.grid)
    else:
        mode = mode or ("autotune" if name.startswith("batch") else "learned")
        plan = spb.Plan(A, mode, nrhs=k)
    build_ms = (time.perf_counter() - t0) * 1e3
    info = plan.info()
    fb_main, fb = plan.format_bytes()
    log(f"plan: {info}  format bytes {fb} (main kernel {fb_main}), CSR bytes {B}")

    for _ in range(warmup):
        plan.spmv_dev(dx.ptr, dy.ptr)
    spb.L.check(lib.vnec_sync_ext(), "sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.spmv_dev(dx.ptr, dy.ptr)
    spb.L.check(lib.vnec_sync_ext(), "sync")
    wall = time.perf_counter() - t0
    ms_step = wall / steps * 1e3
    y_gpu = dy.to_array(np.float64, m.nrows * k)

    k_ms, fix_ms = plan.time_kernels(dx.ptr, dy.ptr, max(steps // 2, 20))
    # hipGraph replay of >= 100 SpMVs per repetition, median of 5 repetitions
    graph_reps = sorted(plan.time(dx.ptr, dy.ptr, max(steps, 100), graph=True) for _ in range(5))
    graph_ms = graph_reps[2]
    log(f"{name}: wall {ms_step * 1e3:.1f} us/step  kernel {k_ms * 1e3:.1f} us  fixup "
        f"{fix_ms * 1e3:.1f} us  graph {graph_ms * 1e3:.1f} us")
    traffic = stored_traffic(name, info)
    rec = {
        "workload": desc, "nrows": m.nrows, "nnz": m.nnz, "nrhs": k,
        "plan": {"variant": info.variant, "ipt": info.ipt, "grid": info.grid, "tiles": info.ntiles,
                 "selection": selection_label(spb, A, None if variant else mode, info, k)},
        "format_bytes": fb, "csr_bytes": B,
        "ms_per_step": round(ms_step, 5),
        "GBps": round(fb / (ms_step / 1e3) / 1e9, 1),
        "csr_equivalent_GBps": round(B / (ms_step / 1e3) / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(fb_main / (k_ms / 1e3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(fb_main / (k_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": kernel_name(info, k), "kernel_us": round(k_ms * 1e3, 2),
                     "fixup_us": round(fix_ms * 1e3, 2),
                     "algorithmic_bytes_per_launch": fb_main,
                     "csr_equivalent_GBps": round(B / (k_ms / 1e3) / 1e9, 1),
                     "counter_GBps": (round(traffic / (k_ms / 1e3) / 1e9, 1) if traffic else None)},
        "graph_spmv_us": round(graph_ms * 1e3, 2),
        "graph_spmv_us_5reps": [round(t * 1e3, 2) for t in graph_reps],
        "plan_build_ms": round(build_ms, 1),
    }
    if flush:
        # the matrix fits the 256 MB Infinity Cache: also time with it flushed
        fl = plan.time(dx.ptr, dy.ptr, 50, graph=False, flush_bytes=R18_FLUSH)
        rec["flushed_spmv_us"] = round(fl * 1e3, 2)
        rec["flushed_GBps"] = round(fb / (fl / 1e3) / 1e9, 1)
        rec["flushed_frac"] = round(fb / (fl / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    if k == 1:
        # what this matrix's traffic costs with no row work: its values and
        # columns streamed, then with one x gather per nonzero
        st, ga = C.c_double(), C.c_double()
        spb.L.check(lib.vnec_gather_ceiling_ext(A.h, dx.ptr, 50, C.byref(st), C.byref(ga)),
                    "gather ceiling")
        rec["gather_ceiling"] = {
            "stream_only_us": round(st.value * 1e3, 2), "stream_gather_us": round(ga.value * 1e3, 2),
            "stream_gather_csr_GBps": round(B / (ga.value / 1e3) / 1e9, 1),
            "kernel_pct_of_stream_gather": round(100.0 * ga.value / k_ms, 1)}
    return rec, m, k, y_gpu, A, dx, dy, plan


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="stencil",
                    choices=["stencil", "rmat18", "skew", "batch", "batch8"])
    ap.add_argument("--variant", default=None, help="force a variant (S, L, E, D, C, G); default selector")
    ap.add_argument("--select", default=None, choices=["learned", "autotune", "rule", "reference"],
                    help="selection mode when no variant is forced (default: autotune for the "
                         "batched workloads, whose tile shapes come from the autotuner; learned "
                         "otherwise)")
    ap.add_argument("--ipt", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0, help="with --variant: 0 persistent, -1 one WG per tile")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-rmat", action="store_true",
                    help="with the default stencil workload: skip the R-MAT 18 record")
    ap.add_argument("--selector-matrices", type=int, default=25,
                    help="held-out matrices for the learned-vs-plain comparison (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(the SpMV would shard by merge-path diagonal ranges, each GPU taking a "
              "contiguous range of rows and exchanging one carry per boundary)",
              file=sys.stderr)
        sys.exit(2)

    from vnec_amd import spb

    if spb.device_count() < 1:
        raise SystemExit("no GPU visible")
    lib = spb.L.load()
    rec, m, k, y_gpu, A, dx, dy, plan = measure(spb, lib, args.workload, args, args.variant,
                                                args.select, flush=args.workload == "rmat18")
    B = rec["csr_bytes"]

    # host vectors (the reference API's semantics: x uploaded, y downloaded per call)
    xh, yh = spb.Vector.from_array(m.x), spb.Vector(m.nrows * k)
    plan.spmv(yh, xh)
    reps_h = 10
    t0 = time.perf_counter()
    for _ in range(reps_h):
        plan.spmv(yh, xh)
    host_ms = (time.perf_counter() - t0) / reps_h * 1e3
    xh.free()
    yh.free()

    variants = {}
    variant_y = {}
    if args.workload == "stencil" and not args.no_variants:
        for name, var, ipt, enc in (("G_fp64", "G", 8, "fp64"), ("G_hilo", "G", 8, "hilo"),
                                    ("G_hi", "G", 8, "hi"), ("S3_fp64", "S", 3, "fp64"),
                                    ("S5_fp64", "S", 5, "fp64"), ("S3_hilo", "S", 3, "hilo"),
                                    ("S3_hi", "S", 3, "hi"), ("L7_fp64", "L", 7, "fp64"),
                                    ("E5_fp64", "E", 5, "fp64")):
            p = spb.Plan(A, var, ipt, encoding=enc)
            p.spmv_dev(dx.ptr, dy.ptr)
            spb.L.check(lib.vnec_sync_ext(), "sync")
            variant_y[name] = dy.to_array(np.float64, m.nrows)
            km, _ = p.time_kernels(dx.ptr, dy.ptr, 50)
            gm = p.time(dx.ptr, dy.ptr, 100, graph=True)
            fbm, fbv = p.format_bytes()
            bv = m.algorithmic_bytes(4 if enc == "hi" else 8)
            variants[name] = {"kernel_us": round(km * 1e3, 2), "spmv_us": round(gm * 1e3, 2),
                              "format_bytes": fbv, "GBps": round(fbv / (gm / 1e3) / 1e9, 1),
                              "frac": round(fbm / (km / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                              "csr_bytes": bv, "csr_equivalent_GBps": round(bv / (gm / 1e3) / 1e9, 1)}
            p.free()

    cp, rd = C.c_double(), C.c_double()  # measured streaming ceilings
    spb.L.check(lib.vnec_copy_bandwidth_ext(1 << 30, 10, C.byref(cp)), "copy bandwidth")
    spb.L.check(lib.vnec_read_bandwidth_ext(1 << 30, 10, C.byref(rd)), "read bandwidth")

    sel = None
    if args.selector_matrices > 0:
        log("selector comparison ...")
        sel = selector_eval(spb, args.selector_matrices)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        with _QuietStdout():
            cpu = cpu_baseline(m, k, y_gpu, variant_y)
        for name, err in cpu.pop("variant_errors").items():
            variants[name]["max_rel_row_err_vs_reference"] = err
    plan.free()

    workloads = {}
    if args.workload == "stencil" and not args.no_rmat:
        # the second configuration of the HBM target, measured the same way
        r, m2, k2, y2, A2, dx2, dy2, p2 = measure(spb, lib, "rmat18", args, args.variant, args.select,
                                                  flush=True)
        p2.free()
        if not args.no_cpu:
            log("cpu baseline (rmat18) ...")
            with _QuietStdout():
                c2 = cpu_baseline(m2, k2, y2, {})
            c2.pop("variant_errors")
            r["cpu_baseline"] = c2
        workloads["rmat18"] = r

    ms_step = rec["ms_per_step"]
    line = {
        "metric": METRIC, "value": rec["GBps"], "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": rec["workload"], "nrows": m.nrows, "nnz": m.nnz, "nrhs": k,
                   "format_bytes": rec["format_bytes"], "csr_bytes": B,
                   "variant": rec["plan"]["variant"], "ipt": rec["plan"]["ipt"],
                   "grid": rec["plan"]["grid"], "tiles": rec["plan"]["tiles"],
                   "selection": rec["plan"]["selection"], "parallelism": "single GPU"},
        "pct_hbm_peak": round(100.0 * rec["GBps"] / HBM_PEAK_GBS, 2),
        "csr_equivalent_GBps": rec["csr_equivalent_GBps"],
        "gflops": round(2.0 * m.nnz * k / (ms_step / 1e3) / 1e9, 1),
        "roofline": rec["roofline"],
        "graph_spmv_us": rec["graph_spmv_us"],
        "graph_spmv_us_5reps": rec["graph_spmv_us_5reps"],
        "host_vectors_spmv_us": round(host_ms * 1e3, 1),
        "host_vectors_GBps": round(rec["format_bytes"] / (host_ms / 1e3) / 1e9, 1),
        "copy_peak_gbs": round(cp.value, 1),
        "read_peak_gbs": round(rd.value, 1),
        "pct_of_read_peak": round(100.0 * rec["roofline"]["achieved"] / rd.value, 2),
        "gather_ceiling": rec.get("gather_ceiling"),
        "plan_build_ms": rec["plan_build_ms"],
        **{key: rec[key] for key in ("flushed_spmv_us", "flushed_GBps", "flushed_frac") if key in rec},
        "variants": variants,
        "workloads": workloads,
        "selector": sel,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
