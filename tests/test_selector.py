"""CPU checks of the selector's checkers and host logic.

What pins each part:
  * integer statistics (oracle/selector_oracle.c): hand-computed values on
    a 4-row matrix, closed forms for banded and block-diagonal matrices,
    brute force (pairwise differences, per-tile sets) on random matrices;
  * the FP64 MLP backward pass (oracle/mlp_oracle.py): central finite
    differences, with a planted sign error that the check must catch;
  * the library's host feature routine (vnec_features_from_ints_ext, C)
    against the FP64 restatement;
  * the committed dataset and model (when present): labels are the argmin
    of the recorded times, the candidate masks match the statistics, and
    the compiled-in model's held-out regret, evaluated by the FP64 oracle,
    is within 10 % of the exhaustive autotuner.
"""
import os

import numpy as np
import pytest

import mlp_oracle as MO
import oracle as O
from vnec_amd import _lib as L
from vnec_amd import matgen as g
from vnec_amd import spb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATASET = os.path.join(ROOT, "vnec_amd", "selector", "dataset.npz")


def _fi(m):
    return O.feature_ints(m.ptr, m.col, m.ncols)


# ------------------------------------------------------------ statistics
def test_feature_ints_hand_example():
    ptr = np.array([0, 2, 2, 5, 6])
    col = np.array([0, 1, 0, 2, 3, 3], np.int32)
    v = O.feature_ints(ptr, col, 4)
    assert list(v[:3]) == [4, 4, 6]
    assert v[MO.F_SUM_LEN2] == 4 + 0 + 9 + 1 and v[MO.F_SUM_LEN3] == 8 + 27 + 1
    assert v[MO.F_MAX_LEN] == 3 and v[MO.F_EMPTY] == 1
    assert list(v[MO.F_HIST:MO.F_HIST + 3]) == [1, 2, 0]
    assert v[MO.F_GINI] == 10                      # |2-0|+|2-3|+|2-1|+|0-3|+|0-1|+|3-1|
    assert v[MO.F_SUM_DIST] == 4 and v[MO.F_MAX_DIST] == 2 and v[MO.F_NEAR_DIAG] == 6
    assert v[MO.F_SLICE_SLOTS] == 64 * 3
    assert v[MO.F_NEC_SMALL] == 4 and v[MO.F_RUNS_SMALL] == 1


def test_feature_ints_banded_closed_form():
    n, hb = 1000, 3
    v = _fi(g.banded(n, hb, seed=1))
    nnz = n * (2 * hb + 1) - hb * (hb + 1)
    assert v[MO.F_NNZ] == nnz and v[MO.F_MAX_LEN] == 2 * hb + 1
    assert v[MO.F_MAX_DIST] == hb and v[MO.F_NEAR_DIAG] == nnz
    # interior rows contribute 2 (1 + .. + hb); edge rows lose their outer band
    full = 2 * sum(range(1, hb + 1))
    edge = sum(sum(range(1, hb + 1)) + sum(range(1, r + 1)) for r in range(hb))
    assert v[MO.F_SUM_DIST] == (n - 2 * hb) * full + 2 * edge
    # every window is one run of consecutive columns
    assert v[MO.F_RUNS_SMALL] == int(np.ceil((n + nnz) / 4096))


def test_feature_ints_blockdiag_closed_form():
    n, bs = 1024, 16
    v = _fi(g.blockdiag(n, bs, seed=2))
    assert v[MO.F_NNZ] == n * bs and v[MO.F_SUM_LEN2] == n * bs * bs
    assert v[MO.F_HIST + 4] == n and v[MO.F_GINI] == 0
    assert v[MO.F_MAX_DIST] == bs - 1
    assert v[MO.F_SLICE_SLOTS] == n * bs               # no padding


@pytest.mark.parametrize("seed", [3, 4])
def test_feature_ints_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = 700
    lens = rng.integers(0, 40, n)
    lens[5] = 70000                                     # clipped at 65535 for the Gini sum
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ncols = 90000
    col = np.concatenate([np.sort(rng.choice(ncols, k, replace=False)) for k in lens]).astype(np.int32)
    v = O.feature_ints(ptr, col, ncols)
    cl = np.minimum(lens, 65535).astype(np.int64)
    assert int(v[MO.F_GINI]) == int(np.abs(cl[:, None] - cl[None, :]).sum() // 2)
    rows = np.repeat(np.arange(n), lens)
    d = np.abs(col.astype(np.int64) - rows * ncols // n)
    assert int(v[MO.F_SUM_DIST]) == int(d.sum()) and int(v[MO.F_MAX_DIST]) == int(d.max())
    for items, ki, kr in ((4096, MO.F_NEC_SMALL, MO.F_RUNS_SMALL),
                          (16384, MO.F_NEC_LARGE, MO.F_RUNS_LARGE)):
        P = -(-(n + len(col)) // items)
        s, e = O.partition(ptr, P)
        nec = runs = 0
        for t in range(P):
            u = np.unique(col[s[t, 1]:e[t, 1]])
            nec += len(u)
            runs += int((np.diff(u) != 1).sum()) + (len(u) > 0)
        assert int(v[ki]) == nec and int(v[kr]) == runs
    slots = sum(64 * lens[i:i + 64].max() for i in range(0, n, 64))
    assert int(v[MO.F_SLICE_SLOTS]) == slots


def test_features_from_ints_host_matches_fp64():
    for m in (g.rmat(11, 8, seed=5), g.powerlaw(3000, 1.8, 2, 2000, seed=6),
              g.banded(2000, 9, seed=7)):
        v = _fi(m)
        f = spb.features_from_ints(v)
        ref = MO.features_from_ints(v)
        np.testing.assert_allclose(f, ref.astype(np.float32), rtol=2e-6, atol=1e-7)


def test_candidate_mask_from_ints():
    n = 640
    lens = np.full(n, 4)
    lens[0] = 4 * 64 * 3                                 # padding just above 3 in slice 0
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate([np.arange(k) for k in lens]).astype(np.int32)
    v = O.feature_ints(ptr, col, 1000)
    cand = spb.candidates()
    m = spb.mask_from_ints(v)
    pad = v[MO.F_SLICE_SLOTS] / v[MO.F_NNZ]
    for c, t in enumerate(cand):
        assert bool((m >> c) & 1) == (t.variant != "S" or pad <= 3.0)


# ------------------------------------------------------------ MLP oracle
def _toy(seed, B=40):
    rng = np.random.default_rng(seed)
    p = MO.init_params(seed).astype(np.float64)
    p[MO.NF * MO.NH:MO.NF * MO.NH + MO.NH] = rng.normal(0, 0.1, MO.NH)   # non-zero b1
    X = rng.normal(size=(B, MO.NF))
    T = rng.normal(size=(B, MO.NC))
    M = (rng.random((B, MO.NC)) < 0.8).astype(np.float64)
    return p, X, T, M


def test_mlp_oracle_backward_matches_finite_differences():
    p, X, T, M = _toy(1)
    loss, grad = MO.loss_grad(p, X, T, M)
    assert np.isclose(loss, MO.loss_only(p, X, T, M))
    rng = np.random.default_rng(2)
    idx = rng.choice(MO.NPARAMS, 120, replace=False)
    fd = MO.fd_grad(p, X, T, M, idx)
    np.testing.assert_allclose(grad[idx], fd, rtol=1e-5, atol=1e-8)


def test_mlp_oracle_fd_catches_planted_error():
    p, X, T, M = _toy(3)
    _, grad = MO.loss_grad(p, X, T, M)
    bad = grad.copy()
    o = MO.NF * MO.NH + MO.NH                           # W2 block: flip its sign
    bad[o:o + MO.NH * MO.NH] *= -1
    idx = np.arange(o, o + 64)
    fd = MO.fd_grad(p, X, T, M, idx)
    assert not np.allclose(bad[idx], fd, rtol=1e-5, atol=1e-8)


def test_param_layout_matches_library():
    assert MO.NPARAMS == L.MLP_NPARAMS and MO.NC == L.NCAND and MO.NF == L.NFEAT


# ------------------------------------------------------------ dataset and model
def _dataset():
    if not os.path.exists(DATASET):
        pytest.skip("selector dataset not generated yet")
    with np.load(DATASET, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_dataset_labels_and_masks():
    d = _dataset()
    t = d["times"]
    assert t.shape[1] == L.NCAND and len(np.unique(d["index"])) == len(d["index"])
    np.testing.assert_array_equal(d["best"], np.argmin(t, axis=1))
    g_bit = 1 << 21  # pattern slices: gated by the row pattern, not the statistics
    for i in range(len(t)):
        m = int(d["mask"][i])
        assert m | g_bit == spb.mask_from_ints(d["ints"][i]) | g_bit
        finite = np.isfinite(t[i])
        assert not (finite & ~((m >> np.arange(L.NCAND)) & 1).astype(bool)).any()
        assert (t[i][finite] >= t[i][d["best"][i]]).all()
    np.testing.assert_allclose(d["feats"], np.stack([spb.features_from_ints(v) for v in d["ints"]]))


def test_compiled_model_regret_on_held_out():
    d = _dataset()
    trained, n = spb.selector_info()
    if not trained:
        pytest.skip("no trained model compiled in")
    from vnec_amd.selector import model_file
    from vnec_amd.selector import specs
    params, mu, sd = model_file.load()
    held = specs.is_held(d["index"])
    assert held.sum() == 200
    t = d["times"][held]
    ch, _ = MO.select(params, mu, sd, d["feats"][held], d["mask"][held])
    best = t.min(axis=1)
    chosen = t[np.arange(len(t)), ch]
    assert np.isfinite(chosen).all()
    agg = chosen.sum() / best.sum()
    geo = np.exp(np.mean(np.log(chosen / best)))
    ref = t[np.arange(len(t)), d["reference"][held]]
    assert agg <= 1.10 and geo <= 1.10, (agg, geo)
    assert np.exp(np.mean(np.log(ref / best))) > geo     # gains over the reference dispatch
    # held-out stencils: pattern slices (candidate 21) are the labels the model must find
    hx = specs.is_held_extra(d["index"])
    if hx.any():
        te = d["times"][hx]
        che, _ = MO.select(params, mu, sd, d["feats"][hx], d["mask"][hx])
        chosen_e = te[np.arange(len(te)), che]
        assert chosen_e.sum() / te.min(axis=1).sum() <= 1.10


def test_compiled_model_source_matches_npz():
    from vnec_amd.selector import model_file
    if not os.path.exists(model_file.NPZ):
        pytest.skip("no trained model")
    trained, p, mu, sd = model_file.parse_c()
    P, MU, SD = model_file.load()
    assert trained == 1
    np.testing.assert_array_equal(p, P)
    np.testing.assert_array_equal(mu, MU)
    np.testing.assert_array_equal(sd, SD)
    assert spb.selector_info() == (True, 2100)
