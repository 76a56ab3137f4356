"""Variant G (pattern slices) through the C-ABI.

The dominant row pattern and the per-row masks and flags are integer
structures: bit-exact against a numpy restatement of the rule (32 sample
rows at multiplicative-hash positions, the most frequent offset list, the
first on ties; a row fits when each of its columns is row + pattern[l] for
a distinct l, its mask holding those l).  SpMV results meet the FP64 gate
|y - y_ref| <= 1e-12 (|A||x|)_i against the serial oracle, for the rows
that fit and for the rest (run by the sub-plan) alike; with FP32 hi-only
slots the bound is 2^-23 (|A||x|)_i.
"""
import numpy as np
import pytest

import oracle as O
from vnec_amd import matgen as g
from vnec_amd import spb

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if spb.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    spb.init()


def np_pattern(ptr, col):
    n = len(ptr) - 1
    samples = []
    for s in range(32):
        i = ((((s + 1) * 2654435761) % 2 ** 32) * n) >> 32
        b, e = ptr[i], ptr[i + 1]
        samples.append(tuple((col[b:e].astype(np.int64) - i).tolist()) if 1 <= e - b <= 32 else None)
    best, best_n = None, 0
    for cand in samples:
        if cand is None:
            continue
        c = sum(1 for t in samples if t == cand)
        if c > best_n:
            best, best_n = cand, c
    return np.array(best if best else [], np.int64)


def np_match(ptr, col, pat):
    """(masks, flags) by a per-row Python walk (small matrices)."""
    n = len(ptr) - 1
    pos = {int(p): l for l, p in enumerate(pat)}
    masks = np.zeros(n, np.uint32)
    flags = np.zeros(n, np.int32)
    for i in range(n):
        m, last = 0, -1
        for c in col[ptr[i]:ptr[i + 1]]:
            l = pos.get(int(c) - i, -1)
            if l <= last:
                flags[i] = 1
                m = 0
                break
            m |= 1 << l
            last = l
        masks[i] = m
    return masks, flags


def planted(n=20000, frac_irregular=0.3, seed=3):
    """Rows follow offsets (-7, -1, 0, 2, 9) except a random 30 %, which get
    random columns; rows near the edges lose out-of-range entries."""
    rng = np.random.default_rng(seed)
    pat = np.array([-7, -1, 0, 2, 9])
    cols = []
    for i in range(n):
        if rng.random() < frac_irregular:
            c = np.unique(rng.integers(0, n, rng.integers(0, 12)))
        else:
            c = i + pat
            c = c[(c >= 0) & (c < n)]
        cols.append(np.sort(c))
    ptr = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64)
    col = np.concatenate(cols).astype(np.int32)
    return g.Csr(n, n, ptr, col, rng.uniform(-1, 1, len(col)), rng.uniform(-1, 1, n), name="planted")


def ring(n=50000, seed=4):
    rng = np.random.default_rng(seed)
    off = np.array([-3, -1, 0, 1, 3])
    col = np.sort((np.arange(n)[:, None] + off[None, :]) % n, axis=1).ravel().astype(np.int32)
    ptr = np.arange(0, 5 * n + 1, 5, dtype=np.int64)
    return g.Csr(n, n, ptr, col, rng.uniform(-1, 1, len(col)), rng.uniform(-1, 1, n), name="ring")


CASES = [("stencil20", g.stencil27(20, seed=5)), ("banded", g.banded(6000, 9, seed=6)),
         ("planted", planted()), ("ring", ring()), ("rmat12", g.rmat(12, 16, seed=7))]


@pytest.mark.parametrize("name,m", CASES, ids=[c[0] for c in CASES])
def test_pattern_and_flags_bit_exact(name, m):
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    pat, nreg = spb.pattern(A)
    ref = np_pattern(m.ptr, m.col)
    assert np.array_equal(pat, ref), (pat, ref)
    if len(pat):
        masks, flags = spb.pattern_match(A, pat)
        ref_masks, ref_flags = np_match(m.ptr, m.col, ref)
        assert np.array_equal(flags, ref_flags)
        assert np.array_equal(masks, ref_masks)
        assert nreg == m.nrows - int(ref_flags.sum())


@pytest.mark.parametrize("name,m", CASES, ids=[c[0] for c in CASES])
def test_g_plan_matches_oracle(name, m):
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    print(f"case {name} variant G", flush=True)
    p = spb.Plan(A, "G")
    x = spb.Vector.from_array(m.x)
    y = spb.Vector(m.nrows)
    p.spmv(y, x)
    yy = y.to_array()
    y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ok, worst = O.check_rows(yy, y0, ax, TOL)
    assert ok, worst
    p.spmv(y, x)  # repeated runs give the same bits
    assert np.array_equal(y.to_array(), yy)
    p.free()


def test_g_known_answer_is_exact():
    """a_ij = i, x_j = j on a stencil: every partial sum is an integer below
    2^53, so any summation order gives the exact y."""
    m = g.stencil27(24, seed=8)
    val = np.repeat(np.arange(m.nrows, dtype=np.float64), np.diff(m.ptr))
    x = np.arange(m.ncols, dtype=np.float64)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, val)
    p = spb.Plan(A, "G")
    y = spb.Vector(m.nrows)
    p.spmv(y, spb.Vector.from_array(x))
    assert np.array_equal(y.to_array(), O.spmv_serial(m.ptr, m.col, val, x))


def test_learned_selector_picks_g_on_stencils_and_rule_mode_keeps_its_rule():
    """Pattern slices are candidate 21 of the learned selector: the MLP picks
    them on a stencil; on R-MAT they do not apply (no row pattern covers
    half of the rows); the rule mode's structural rule still picks them on
    large stencils only."""
    big = g.stencil27(64, seed=9)  # 262,144 rows, 6.9M nonzeros; every row fits the pattern
    A = spb.Matrix.from_csr(big.nrows, big.ncols, big.ptr, big.col, big.val)
    assert (spb.candidate_mask(A) >> 21) & 1
    assert spb.candidates()[spb.select(A, "learned")].variant == "G"
    p = spb.Plan(A, "learned")
    assert p.info().variant == "G"
    y = spb.Vector(big.nrows)
    p.spmv(y, spb.Vector.from_array(big.x))
    y0 = O.spmv_serial(big.ptr, big.col, big.val, big.x)
    ax, _ = O.spmv_abs(big.ptr, big.col, big.val, big.x)
    assert O.check_rows(y.to_array(), y0, ax, TOL)[0]
    p.free()
    assert spb.Plan(A, "rule").info().variant == "G"
    small = g.stencil27(16, seed=10)  # below 2^20 nonzeros: the rule leaves it to slices / L
    B = spb.Matrix.from_csr(small.nrows, small.ncols, small.ptr, small.col, small.val)
    assert spb.Plan(B, "rule").info().variant != "G"
    R = spb.Matrix.from_csr(*(lambda r: (r.nrows, r.ncols, r.ptr, r.col, r.val))(g.rmat(17, 16, seed=11)))
    assert not (spb.candidate_mask(R) >> 21) & 1
    assert spb.Plan(R, "learned").info().variant != "G"
    assert spb.Plan(R, "rule").info().variant != "G"


@pytest.mark.parametrize("enc,tol", [("hilo", TOL), ("hi", 2.0 ** -23)])
@pytest.mark.parametrize("name,m", CASES, ids=[c[0] for c in CASES])
def test_g_value_encodings(name, m, enc, tol):
    """FP32 hi + lo slots (a = hi + lo exactly to 2^-48 |a|) meet the FP64
    gate; hi-only slots err by at most 2^-24 |a| per term, i.e. within
    2^-23 (|A||x|)_i, and must really lose bits.  The irregular rows' sub-plan
    uses the same encoding."""
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    print(f"case {name} variant G encoding {enc}", flush=True)
    p = spb.Plan(A, "G", encoding=enc)
    assert p.info().encoding == enc
    y = spb.Vector(m.nrows)
    p.spmv(y, spb.Vector.from_array(m.x))
    y0 = O.spmv_serial(m.ptr, m.col, m.val, m.x)
    ax, _ = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    ok, worst = O.check_rows(y.to_array(), y0, ax, tol)
    assert ok, worst
    if enc == "hi":
        assert worst > 1e-12
    p.free()


@pytest.mark.parametrize("name,m", CASES, ids=[c[0] for c in CASES])
def test_g_fp32_api(name, m):
    """The FP32 API's arithmetic (FP32 values, x, y and sums) in pattern
    slices, irregular rows through an FP32 sub-plan: the FP32 gate
    |y - y_ref| <= max(1e-4, 2 (len_i + 1) 2^-24) (|A||x|)_i against the FP32
    serial oracle, and bit-identical repeats."""
    v32, x32 = m.val.astype(np.float32), m.x.astype(np.float32)
    A32 = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, v32, dtype=np.float32)
    print(f"case {name} variant G FP32", flush=True)
    p = spb.Plan(A32, "G")
    assert p.info().encoding == "fp32"
    xv = spb.Vector.from_array(x32, np.float32)
    y = spb.Vector(m.nrows, np.float32)
    p.spmv(y, xv)
    yy = y.to_array()
    ax, _ = O.spmv_abs(m.ptr, m.col, v32.astype(np.float64), x32.astype(np.float64))
    ok, worst = O.check_rows_f32(yy, O.spmv_serial_f32(m.ptr, m.col, v32, x32), m.ptr, ax)
    assert ok, worst
    p.spmv(y, xv)
    assert np.array_equal(y.to_array(), yy)
    p.free()


def test_g_refuses_what_it_cannot_store():
    E = spb.Matrix.from_csr(4, 4, np.zeros(5, np.int64), np.zeros(0, np.int32), np.zeros(0))
    with pytest.raises(RuntimeError, match="no row pattern"):
        spb.Plan(E, "G")


@pytest.mark.parametrize("enc", ["hilo", "hi"])
def test_g_encoded_sub_plan_for_every_selector_pick(enc):
    """The irregular rows' sub-plan takes the selector's pick in the G plan's
    encoding; picks that store FP64 values only (C, D, H) run as the
    merge-path kernel instead.  A planted matrix whose irregular rows are
    R-MAT-like exercises it."""
    m = planted(n=60000, frac_irregular=0.4, seed=21)
    A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, m.val)
    p = spb.Plan(A, "G", encoding=enc)
    y = spb.Vector(m.nrows)
    p.spmv(y, spb.Vector.from_array(m.x))
    ax, yl = O.spmv_abs(m.ptr, m.col, m.val, m.x)
    assert O.check_rows(y.to_array(), yl, ax, TOL if enc == "hilo" else 2.0 ** -23)[0]
    p.free()


@pytest.mark.parametrize("name,m", [CASES[0], CASES[2]], ids=[CASES[0][0], CASES[2][0]])
def test_g_x_at_any_alignment(name, m):
    """The LDS window is staged with 16-byte loads aligned to x's own
    address: x and y at every 4- or 8-byte offset from a 16-byte boundary
    (device pointers from the caller) give the aligned call's bits, FP64 and
    the FP32 API alike, and meet the gate."""
    L = spb.L.load()
    for dtype, offs in ((np.float64, (0, 8)), (np.float32, (0, 4, 8, 12))):
        val, x = m.val.astype(dtype), m.x.astype(dtype)
        A = spb.Matrix.from_csr(m.nrows, m.ncols, m.ptr, m.col, val, dtype=dtype)
        p = spb.Plan(A, "G")
        es = np.dtype(dtype).itemsize
        dx = spb.DeviceArray(es * m.ncols + 16)
        dy = spb.DeviceArray(es * m.nrows + 16)
        outs = []
        for off in offs:
            print(f"case {name} variant G {np.dtype(dtype).name} offset {off}", flush=True)
            spb.L.check(L.vnec_memcpy_h2d_ext(dx.ptr + off, x.ctypes.data, x.nbytes), "h2d")
            p.spmv_dev(dx.ptr + off, dy.ptr + off)
            y = np.empty(m.nrows, dtype)
            spb.L.check(L.vnec_memcpy_d2h_ext(y.ctypes.data, dy.ptr + off, y.nbytes), "d2h")
            outs.append(y)
        for y in outs[1:]:
            assert np.array_equal(y, outs[0])
        v64, x64 = val.astype(np.float64), x.astype(np.float64)
        ax, _ = O.spmv_abs(m.ptr, m.col, v64, x64)
        if dtype == np.float64:
            assert O.check_rows(outs[0], O.spmv_serial(m.ptr, m.col, v64, x64), ax, TOL)[0]
        else:
            assert O.check_rows_f32(outs[0], O.spmv_serial_f32(m.ptr, m.col, val, x), m.ptr, ax)[0]
        dx.free()
        dy.free()
        p.free()
        A.free()
