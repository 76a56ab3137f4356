"""Train the selector MLP on the GPU and evaluate it on the held-out set.

Inputs: the labelled dataset (gen_dataset.py).  Targets are
log2(t_c / t_best), clipped at 3, for every candidate c the autotuner ran
(mask 1), none for those that do not apply (mask 0).  Features are
standardised with the training split's mean and standard deviation.

Training runs the library's HIP kernels (vnec_mlp_train_ext: forward,
backward, deterministic chunk reduction, Adam).  Evaluation runs the HIP
inference kernel (vnec_mlp_select_ext) on the held-out matrices and
reports the regret of its picks against the exhaustive autotuner, beside
the rule selector, the reference dispatch and the best single fixed
candidate.  Writes model.npz and csrc/host/spb_selector_model.c.

    python -m vnec_amd.selector.train --data vnec_amd/selector/dataset.npz
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .. import _lib as L
from .. import spb
from . import model_file, specs

CLIP = 3.0


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    nf, nh, nc = L.NFEAT, L.MLP_HIDDEN, L.NCAND
    parts = []
    for fan_in, shape in ((nf, (nf, nh)), (0, (nh,)), (nh, (nh, nh)), (0, (nh,)),
                          (nh, (nh, nc)), (0, (nc,))):
        if fan_in:
            lim = np.sqrt(6.0 / fan_in)
            parts.append(rng.uniform(-lim, lim, shape).ravel())
        else:
            parts.append(np.zeros(shape))
    return np.concatenate(parts).astype(np.float32)


def targets(times: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    fin = np.isfinite(times)
    best = times.min(axis=1, keepdims=True)
    T = np.where(fin, np.minimum(np.log2(np.where(fin, times, 1.0) / best), CLIP), 0.0)
    return T.astype(np.float32), fin.astype(np.float32)


def regret(times: np.ndarray, choice: np.ndarray) -> dict:
    best = times.min(axis=1)
    c = times[np.arange(len(times)), choice]
    r = c / best
    return {"aggregate": float(c.sum() / best.sum()), "geomean": float(np.exp(np.log(r).mean())),
            "mean": float(r.mean()), "max": float(r.max()), "within_10pct": float((r <= 1.10).mean())}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--data", default=model_file.HERE + "/dataset.npz")
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--wd", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--npz", default=model_file.NPZ)
    ap.add_argument("--c-src", default=model_file.C_SRC)
    args = ap.parse_args(argv)
    with np.load(args.data, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    tr = specs.is_train(d["index"])
    held = specs.is_held(d["index"])
    held_extra = specs.is_held_extra(d["index"])
    F = d["feats"].astype(np.float32)
    mu = F[tr].mean(axis=0)
    sd = F[tr].std(axis=0)
    sd[sd < 1e-6] = 1.0
    Z = ((F - mu) / sd).astype(np.float32)
    T, M = targets(d["times"])
    spb.init()
    p0 = init_params(args.seed)
    t0 = time.time()
    params, hist = spb.mlp_train(p0, Z[tr], T[tr], M[tr], args.epochs, args.batch, args.lr,
                                 args.wd, args.seed)
    train_s = time.time() - t0
    for e in range(0, args.epochs, max(1, args.epochs // 10)):
        print(f"epoch {e + 1:4d} loss {hist[e]:.5f}", flush=True)
    print(f"epoch {args.epochs:4d} loss {hist[-1]:.5f}  ({train_s:.2f} s on the GPU)", flush=True)
    t = d["times"]
    out = {}
    for name, m in (("train", tr), ("held_out", held), ("held_out_stencils", held_extra)):
        if not m.any():
            continue
        ch = spb.mlp_select(params, mu, sd, F[m], d["mask"][m].astype(np.uint32))
        fixed = int(np.argmin([np.exp(np.log(np.where(np.isfinite(t[tr][:, c]), t[tr][:, c], 1e30)
                                            / t[tr].min(1)).mean()) for c in range(L.NCAND)]))
        out[name] = {"n": int(m.sum()), "learned": regret(t[m], ch),
                     "rule": regret(t[m], d["rule"][m]),
                     "reference": regret(t[m], d["reference"][m]),
                     f"fixed_candidate_{fixed}": regret(t[m], np.full(int(m.sum()), fixed))}
    summary = {"epochs": args.epochs, "batch": args.batch, "lr": args.lr, "weight_decay": args.wd,
               "seed": args.seed, "clip": CLIP, "final_loss": float(hist[-1]),
               "train_seconds": train_s, **out}
    print(json.dumps(summary, indent=1, sort_keys=True), flush=True)
    model_file.write(params, mu, sd, int(tr.sum()), summary, args.npz, args.c_src)
    print(f"wrote {args.npz} and {args.c_src}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
