"""Seeded specifications of the selector's training and held-out matrices.

Matrix i (0 <= i < 2200; 0..1999 train, 2000..2199 held out) comes from
family i % 5 with parameters drawn from numpy's PCG64 seeded with
(20250, i):

  rmat       scale 14-20, edge factor 4-32 (edges capped at 2^24),
             a in [0.45, 0.65], b = c, d = 1 - a - 2b >= 0.02
  banded     n log-uniform in [2^13, 2^20], half-bandwidth 1-64
  blockdiag  n log-uniform, block size 8-512 (log-uniform)
  uniform    n log-uniform, 2-128 entries per row (log-uniform)
  powerlaw   n log-uniform, row lengths xmin * U^(-1/(alpha-1)),
             alpha in [1.5, 3], xmin 1-4, capped at min(n, 32768)

Every family caps its expected entry count at 2^24 by shrinking n.
Values and x are FP64 uniform[-1, 1] from the generator's own seed.

Beyond these 2200, matrices 2200..2319 (2200..2299 train, 2300..2319 held
out) are grid stencils, a structure none of the five families has (rows
holding a few short runs of columns at fixed offsets, the layout pattern
slices serve):

  stencil    2-D 5- or 9-point on nx x ny, or 3-D 7- or 27-point on
             nx x ny x nz, grid points log-uniform in [2^13, 2^20] (at most
             2^24 entries), sides within a factor 2 of each other

Stencil values: diagonal (points - 1) + U[0, 1), off-diagonals U[-1, 0).
"""
from __future__ import annotations

import json

import numpy as np

from .. import matgen

N_TRAIN = 2000
N_HELD = 200
N_TOTAL = N_TRAIN + N_HELD
N_EXTRA_TRAIN = 100  # stencils 2200..2299
N_EXTRA_HELD = 20    # stencils 2300..2319
N_ALL = N_TOTAL + N_EXTRA_TRAIN + N_EXTRA_HELD
FAMILIES = ("rmat", "banded", "blockdiag", "uniform", "powerlaw")

SEED_BASE = 20250
MAX_NNZ = 1 << 24


def is_train(i):
    """Training matrices: the five families' first 2000 and the first 100
    stencils (works elementwise on index arrays)."""
    i = np.asarray(i)
    return (i < N_TRAIN) | ((i >= N_TOTAL) & (i < N_TOTAL + N_EXTRA_TRAIN))


def is_held(i):
    """The five families' held-out matrices."""
    i = np.asarray(i)
    return (i >= N_TRAIN) & (i < N_TOTAL)


def is_held_extra(i):
    """Held-out stencils."""
    return np.asarray(i) >= N_TOTAL + N_EXTRA_TRAIN


def _logu(rng, lo: float, hi: float) -> float:
    return float(2.0 ** rng.uniform(np.log2(lo), np.log2(hi)))


def _powerlaw_mean(alpha: float, xmin: int, cap: int) -> float:
    u = 1.0 - np.random.default_rng(7).random(20000)
    return float(np.minimum(np.floor(xmin * u ** (-1.0 / (alpha - 1.0))), cap).mean())


def spec(i: int) -> dict:
    rng = np.random.default_rng([SEED_BASE, i])
    fam = FAMILIES[i % len(FAMILIES)] if i < N_TOTAL else "stencil"
    s: dict = {"index": i, "family": fam, "seed": SEED_BASE * 10000 + i}
    if fam == "stencil":
        pts = int(rng.choice([5, 9, 7, 27]))
        npts = _logu(rng, 1 << 13, 1 << 20)
        npts = min(npts, MAX_NNZ / pts)
        dims = 2 if pts in (5, 9) else 3
        aspect = 2.0 ** rng.uniform(-1, 1, dims)
        aspect /= np.prod(aspect) ** (1.0 / dims)
        side = npts ** (1.0 / dims)
        n = [max(4, int(round(side * a))) for a in aspect]
        while int(np.prod(n)) * pts > MAX_NNZ:
            n[int(np.argmax(n))] -= 1
        s.update(pts=pts, nx=n[0], ny=n[1], nz=n[2] if dims == 3 else 1)
    elif fam == "rmat":
        scale = int(rng.integers(14, 21))
        ef = int(round(_logu(rng, 4, 32)))
        ef = max(1, min(ef, MAX_NNZ >> scale))
        a = float(rng.uniform(0.45, 0.65))
        b = float(rng.uniform(0.05, (1.0 - a - 0.02) / 2))
        s.update(scale=scale, ef=ef, a=a, b=b, c=b)
    else:
        n = int(_logu(rng, 1 << 13, 1 << 20))
        if fam == "banded":
            hb = int(rng.integers(1, 65))
            s.update(n=min(n, MAX_NNZ // (2 * hb + 1)), hb=hb)
        elif fam == "blockdiag":
            bs = int(round(_logu(rng, 8, 512)))
            s.update(n=min(n, MAX_NNZ // bs), bs=bs)
        elif fam == "uniform":
            k = int(round(_logu(rng, 2, 128)))
            s.update(n=min(n, MAX_NNZ // k), k=k)
        else:
            alpha = float(rng.uniform(1.5, 3.0))
            xmin = int(rng.integers(1, 5))
            cap = min(n, 1 << 15)
            mean = max(_powerlaw_mean(alpha, xmin, cap), 1.0)
            n = min(n, int(MAX_NNZ / mean))
            s.update(n=n, alpha=alpha, xmin=xmin, cap=min(n, 1 << 15))
    return s


def generate(s: dict) -> matgen.Csr:
    fam = s["family"]
    if fam == "stencil":
        return matgen.stencil(s["nx"], s["ny"], s["nz"], s["pts"], seed=s["seed"])
    if fam == "rmat":
        return matgen.rmat(s["scale"], s["ef"], s["a"], s["b"], s["c"], seed=s["seed"])
    if fam == "banded":
        return matgen.banded(s["n"], s["hb"], seed=s["seed"])
    if fam == "blockdiag":
        return matgen.blockdiag(s["n"], s["bs"], seed=s["seed"])
    if fam == "uniform":
        return matgen.uniform(s["n"], s["k"], seed=s["seed"])
    return matgen.powerlaw(s["n"], s["alpha"], s["xmin"], s["cap"], seed=s["seed"])


def spec_json(s: dict) -> str:
    return json.dumps(s, sort_keys=True)
