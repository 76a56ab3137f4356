"""Label the selector's seeded matrices with the exhaustive autotuner.

For each matrix index in [start, start + count) (see specs.py): generate it
on host threads (ahead of the GPU), upload it, compute the 32 integer
statistics on the GPU, time every applicable candidate (hipGraph replay,
best of 3 means of `reps` SpMVs) and record the choices of the rule and
reference-dispatch selectors.  Results are written to an .npz file (also
every 25 matrices, so a cut-off run keeps what it measured).

    python -m vnec_amd.selector.gen_dataset --start 0 --count 550 --out ds_0.npz
"""
from __future__ import annotations

import argparse
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .. import _lib as L
from .. import spb
from . import specs


def _save(path: str, rec: dict):
    tmp = path + ".tmp.npz"
    np.savez_compressed(tmp, **{k: np.asarray(v) for k, v in rec.items()})
    os.replace(tmp, path)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--start", type=int, default=0)
    ap.add_argument("--count", type=int, default=specs.N_ALL)
    ap.add_argument("--out", required=True)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--threads", type=int, default=6, help="generator threads")
    args = ap.parse_args(argv)
    idx = list(range(args.start, min(args.start + args.count, specs.N_ALL)))
    spb.init()
    rec: dict[str, list] = {k: [] for k in ("index", "spec", "ints", "feats", "mask", "times",
                                            "best", "rule", "reference", "gen_s", "tune_s")}
    t0 = time.time()
    pool = ThreadPoolExecutor(max_workers=args.threads)
    ahead = args.threads + 2

    def gen(i):
        s = specs.spec(i)
        t = time.time()
        m = specs.generate(s)
        return s, m, time.time() - t

    futs = {i: pool.submit(gen, i) for i in idx[:ahead]}
    for n, i in enumerate(idx):
        s, m, gs = futs.pop(i).result()
        if n + ahead < len(idx):
            j = idx[n + ahead]
            futs[j] = pool.submit(gen, j)
        t = time.time()
        A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
        del m
        ints = spb.feature_ints(A)
        feats = spb.features_from_ints(ints)
        mask = spb.candidate_mask(A)  # the statistics' rules and the row-pattern gate
        times, best = spb.autotune(A, args.reps)
        rule = spb.select(A, "rule")
        ref = spb.select(A, "reference")
        A.free()
        ts = time.time() - t
        for k, v in (("index", i), ("spec", specs.spec_json(s)), ("ints", ints), ("feats", feats),
                     ("mask", mask), ("times", times), ("best", best), ("rule", rule),
                     ("reference", ref), ("gen_s", gs), ("tune_s", ts)):
            rec[k].append(v)
        print(f"[{i}] {s['family']:9s} nnz {int(ints[2]):>9d} best {best:2d} "
              f"{times[best] * 1e3:8.1f} us  rule {times[rule] / times[best]:.2f}x  "
              f"ref {times[ref] / times[best]:.2f}x  gen {gs:.1f}s tune {ts:.2f}s  "
              f"({n + 1}/{len(idx)}, {time.time() - t0:.0f}s)", flush=True)
        if (n + 1) % 25 == 0:
            _save(args.out, rec)
    pool.shutdown()
    _save(args.out, rec)
    print(f"wrote {len(idx)} matrices to {args.out} in {time.time() - t0:.0f}s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
