"""Seeded synthetic matrices for the VNEC SpMV workloads (ctypes over
libvnec_matgen.so, built from csrc/gen/matgen.c).

Each generator returns a :class:`Csr` with int64 ``ptr``, int32 ``col``
(sorted and unique within each row), float64 ``val`` and float64 ``x``.
The draw order is fixed (structure, values in CSR order, x), so a seed
gives identical bytes everywhere.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libvnec_matgen.so")


class _VgCsr(C.Structure):
    _fields_ = [
        ("nrows", C.c_int64), ("ncols", C.c_int64), ("nnz", C.c_int64),
        ("ptr", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int32)),
        ("val", C.POINTER(C.c_double)), ("x", C.POINTER(C.c_double)),
        ("nrhs", C.c_int64), ("nmat", C.c_int64),
        ("row_off", C.POINTER(C.c_int64)), ("col_off", C.POINTER(C.c_int64)),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built; run __graft_entry__.build()")
        lib = C.CDLL(_LIB_PATH)
        P = C.POINTER(_VgCsr)
        for name, args in {
            "vg_rmat": [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64],
            "vg_skew": [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                        C.c_int64, C.c_uint64],
            "vg_stencil27": [C.c_int, C.c_uint64],
            "vg_stencil": [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64],
            "vg_batch": [C.c_int, C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_int64,
                         C.c_uint64],
            "vg_banded": [C.c_int64, C.c_int, C.c_uint64],
            "vg_blockdiag": [C.c_int64, C.c_int, C.c_uint64],
            "vg_uniform": [C.c_int64, C.c_int, C.c_uint64],
            "vg_powerlaw": [C.c_int64, C.c_double, C.c_int, C.c_int64, C.c_uint64],
            "vg_from_pattern": [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_uint64],
        }.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = P
        lib.vg_free.argtypes = [P]
        lib.vg_free.restype = None
        lib.vg_stream.argtypes = [C.c_uint64, C.c_int64, C.c_void_p]
        lib.vg_stream.restype = None
        _lib = lib
    return _lib


@dataclass
class Csr:
    nrows: int
    ncols: int
    ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    x: np.ndarray
    nrhs: int = 1
    row_off: np.ndarray | None = None
    col_off: np.ndarray | None = None
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    def algorithmic_bytes(self, vb: int = 8, k: int = 1) -> int:
        """B = (vb + 4) nnz + 4 (nrows + 1) + 8 k ncols + 8 k nrows."""
        return ((vb + 4) * self.nnz + 4 * (self.nrows + 1) + 8 * k * self.ncols
                + 8 * k * self.nrows)


def _take(p, name="") -> Csr:
    m = p.contents
    n, c, z, k = m.nrows, m.ncols, m.nnz, m.nrhs
    ptr = np.ctypeslib.as_array(m.ptr, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(m.col, shape=(max(z, 1),))[:z].copy()
    val = np.ctypeslib.as_array(m.val, shape=(max(z, 1),))[:z].copy()
    x = np.ctypeslib.as_array(m.x, shape=(max(c * k, 1),))[: c * k].copy()
    ro = co = None
    if m.nmat:
        ro = np.ctypeslib.as_array(m.row_off, shape=(m.nmat + 1,)).copy()
        co = np.ctypeslib.as_array(m.col_off, shape=(m.nmat + 1,)).copy()
    _lib.vg_free(p)
    return Csr(n, c, ptr, col, val, x, k, ro, co, name)


def rmat(scale: int, ef: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1) -> Csr:
    return _take(_load().vg_rmat(scale, ef, a, b, c, seed), f"rmat{scale}")


def skew(scale: int = 22, ef: int = 16, ndense: int = 8, dense_len: int = 500_000, a=0.57,
         b=0.19, c=0.19, seed: int = 1) -> Csr:
    return _take(_load().vg_skew(scale, ef, a, b, c, ndense, dense_len, seed), f"skew{scale}")


def stencil27(n: int = 128, seed: int = 1) -> Csr:
    return _take(_load().vg_stencil27(n, seed), f"stencil27_{n}")


def batch(nmat: int = 512, rmin: int = 1000, rmax: int = 200_000, mu_lo=4.0, mu_hi=64.0,
          nrhs: int = 1, seed: int = 1) -> Csr:
    """Block-diagonal batch; ``col`` holds LOCAL column ids of each member."""
    return _take(_load().vg_batch(nmat, rmin, rmax, mu_lo, mu_hi, nrhs, seed), f"batch{nmat}")


def stencil(nx: int, ny: int, nz: int, pts: int, seed: int = 1) -> Csr:
    """Grid stencil: pts = 5 or 9 on an nx x ny grid (nz = 1), 7 or 27 on
    an nx x ny x nz grid; truncated at the boundary."""
    if pts in (5, 9) and nz != 1 or pts in (7, 27) and nz < 2 or pts not in (5, 9, 7, 27):
        raise ValueError(f"stencil: pts {pts} on a {nx} x {ny} x {nz} grid")
    return _take(_load().vg_stencil(nx, ny, nz, pts, seed), f"stencil{pts}_{nx}x{ny}x{nz}")


def banded(n: int, hb: int, seed: int = 1) -> Csr:
    return _take(_load().vg_banded(n, hb, seed), f"banded{hb}")


def blockdiag(n: int, bs: int, seed: int = 1) -> Csr:
    return _take(_load().vg_blockdiag(n, bs, seed), f"blockdiag{bs}")


def uniform(n: int, k: int, seed: int = 1) -> Csr:
    return _take(_load().vg_uniform(n, k, seed), f"uniform{k}")


def powerlaw(n: int, alpha: float, xmin: int = 2, cap: int | None = None, seed: int = 1) -> Csr:
    cap = n if cap is None else cap
    return _take(_load().vg_powerlaw(n, alpha, xmin, cap, seed), f"powerlaw{alpha:.2f}")


def from_pattern(nrows: int, ncols: int, ptr, col, seed: int = 1) -> Csr:
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    return _take(_load().vg_from_pattern(nrows, ncols, ptr.ctypes.data, col.ctypes.data, seed),
                 "pattern")


def read_mtx(path: str) -> tuple[int, int, np.ndarray, np.ndarray, np.ndarray]:
    """MatrixMarket coordinate file -> (nrows, ncols, ptr, col, val), CSR
    sorted by (row, col); duplicates kept in file order."""
    with open(path) as f:
        lines = [ln for ln in f if ln.strip() and not ln.startswith("%")]
    nr, nc, nz = (int(t) for t in lines[0].split()[:3])
    ent = [ln.split() for ln in lines[1:1 + nz]]
    r = np.array([int(e[0]) - 1 for e in ent], dtype=np.int64)
    c = np.array([int(e[1]) - 1 for e in ent], dtype=np.int64)
    v = np.array([float(e[2]) if len(e) > 2 else 1.0 for e in ent], dtype=np.float64)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    ptr = np.zeros(nr + 1, dtype=np.int64)
    np.add.at(ptr, r + 1, 1)
    ptr = np.cumsum(ptr)
    return nr, nc, ptr, c.astype(np.int32), v


def stream(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    _load().vg_stream(seed, n, out.ctypes.data)
    return out
