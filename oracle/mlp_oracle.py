"""FP64 CPU restatement of the selector: features from integer statistics,
the MLP's forward and backward passes, and masked argmin selection.

TEST INFRASTRUCTURE ONLY: the checker for the library's HIP MLP kernels
(vnec_select.hip) and its host feature routine (vnec_features_from_ints_ext).
The selector has no counterpart in the reference; the definitions restated
here are the library's own (include/vnec_ext.h).  The backward pass is
written out by hand and is itself pinned by central finite differences
(tests/test_selector.py).
"""
from __future__ import annotations

import numpy as np

NF, NH, NC = 32, 64, 23
SHAPES = [("W1", (NF, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NC)), ("b3", (NC,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)

# integer statistic indices (include/vnec_ext.h)
(F_NROWS, F_NCOLS, F_NNZ, F_SUM_LEN2, F_SUM_LEN3, F_MAX_LEN, F_EMPTY, F_HIST) = range(8)
F_GINI = F_HIST + 16
(F_SUM_DIST, F_MAX_DIST, F_NEAR_DIAG, F_SLICE_SLOTS, F_NEC_SMALL, F_RUNS_SMALL, F_NEC_LARGE,
 F_RUNS_LARGE) = range(F_GINI + 1, F_GINI + 9)


def unpack(params) -> dict[str, np.ndarray]:
    p = np.asarray(params, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d: dict[str, np.ndarray]) -> np.ndarray:
    return np.concatenate([np.asarray(d[n], dtype=np.float64).ravel() for n, _ in SHAPES])


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform initial weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shape[0])
            d[name] = rng.uniform(-lim, lim, shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)


def forward(params, X):
    """Returns (outputs [B, NC], cache) in FP64."""
    P = unpack(params)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ P["W1"] + P["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ P["W2"] + P["b2"]
    h2 = np.maximum(a2, 0.0)
    o = h2 @ P["W3"] + P["b3"]
    return o, (X, a1, h1, a2, h2)


def loss_grad(params, X, T, M):
    """Loss sum M (o - T)^2 / sum M and its gradient (flat, FP64)."""
    P = unpack(params)
    o, (X, a1, h1, a2, h2) = forward(params, X)
    T = np.asarray(T, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    den = M.sum()
    diff = o - T
    loss = float((M * diff * diff).sum() / den)
    dO = 2.0 * M * diff / den
    g = {"W3": h2.T @ dO, "b3": dO.sum(0)}
    d2 = (dO @ P["W3"].T) * (a2 > 0)
    g["W2"] = h1.T @ d2
    g["b2"] = d2.sum(0)
    d1 = (d2 @ P["W2"].T) * (a1 > 0)
    g["W1"] = X.T @ d1
    g["b1"] = d1.sum(0)
    return loss, pack(g)


def loss_only(params, X, T, M) -> float:
    o, _ = forward(params, X)
    M = np.asarray(M, dtype=np.float64)
    d = o - np.asarray(T, dtype=np.float64)
    return float((M * d * d).sum() / M.sum())


def fd_grad(params, X, T, M, idx, h: float = 1e-6) -> np.ndarray:
    """Central finite differences of the loss at parameter indices idx."""
    p = np.asarray(params, dtype=np.float64).copy()
    out = np.zeros(len(idx))
    for n, i in enumerate(idx):
        keep = p[i]
        p[i] = keep + h
        lp = loss_only(p, X, T, M)
        p[i] = keep - h
        lm = loss_only(p, X, T, M)
        p[i] = keep
        out[n] = (lp - lm) / (2 * h)
    return out


def features_from_ints(v) -> np.ndarray:
    """The 32 selector inputs from the integer statistics (FP64)."""
    v = np.asarray(v, dtype=np.uint64).astype(np.float64)
    n, m, z = v[F_NROWS], v[F_NCOLS], v[F_NNZ]
    nn, zz, mm = (n if n > 0 else 1.0), (z if z > 0 else 1.0), (m if m > 0 else 1.0)
    mean = z / nn
    ex2, ex3 = v[F_SUM_LEN2] / nn, v[F_SUM_LEN3] / nn
    var = ex2 - mean * mean
    sd = np.sqrt(var) if var > 0 else 0.0
    skew = (ex3 - 3 * mean * var - mean ** 3) / sd ** 3 if sd > 0 else 0.0
    skew = min(max(skew, -50.0), 50.0)
    f = [np.log2(n + 1), np.log2(m + 1), np.log2(z + 1), np.log2(mean + 1),
         sd / mean if mean > 0 else 0.0, skew / 10.0, np.log2(v[F_MAX_LEN] + 1),
         v[F_EMPTY] / nn]
    f += [v[F_HIST + b] / nn for b in range(16)]
    f += [v[F_GINI] / (nn * zz), v[F_SUM_DIST] / zz / mm, v[F_MAX_DIST] / mm,
          v[F_NEAR_DIAG] / zz, v[F_NEC_SMALL] / zz, v[F_RUNS_SMALL] / zz, v[F_NEC_LARGE] / zz,
          np.log2(v[F_SLICE_SLOTS] / zz + 1)]
    return np.array(f, dtype=np.float64)


def select(params, mu, sd, F, masks) -> tuple[np.ndarray, np.ndarray]:
    """(choice per row, FP64 outputs) with the inference kernel's rule:
    argmin over allowed candidates, lowest index on ties, -1 if none."""
    Z = (np.asarray(F, np.float64) - np.asarray(mu, np.float64)) / np.asarray(sd, np.float64)
    o, _ = forward(params, Z)
    masks = np.asarray(masks, dtype=np.uint64)
    allowed = ((masks[:, None] >> np.arange(NC, dtype=np.uint64)[None, :]) & 1).astype(bool)
    v = np.where(allowed, o, np.inf)
    ch = np.argmin(v, axis=1)
    ch[~allowed.any(axis=1)] = -1
    return ch, o
