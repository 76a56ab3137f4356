"""Python mirror of the VNEC path of the OpenSparseBLAS API, over the HIP
C-ABI of libvnec_mi355x.so.

Names follow the reference (SpB_Matrix / SpB_Vector objects, IRD_VNEC,
Build_VNEC_* / SpMV_VNEC_*_FP64 and _FP32, SpB_mxv); errors raise
RuntimeError with the library's reason.  Everything runs on the GPU through
the library.  Objects hold FP64 values unless made with dtype=np.float32
(the reference's FP32 API: FP32 values, x, y and sums).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

SpB_FP32 = 10
SpB_FP64 = 11
SpB_MATRIX_FORMAT = 1
SpB_CSR = 1
VNEC_VAR = {"L": 0, "E": 1, "S": 2, "D": 3, "H": 4, "G": 5, "C": 6}
VNEC_ENC = {"fp64": 0, "hilo": 1, "hi": 2, "fp32": 3}
# Selection modes (negative vnec_tuning.variant values).
SELECT = {"learned": L.SELECT_LEARNED, "autotune": L.SELECT_AUTOTUNE, "rule": L.SELECT_RULE,
          "reference": L.SELECT_REFERENCE}

_init = False


def init():
    global _init
    lib = L.load()
    if not _init:
        lib.SpB_init()
        f = L.SpB_Object_Format_ext()
        f.matrix_format = SpB_CSR
        L.check(lib.SpB_Global_Option_set_ext(SpB_MATRIX_FORMAT, f), "SpB_Global_Option_set_ext")
        _init = True
    return lib


def device_count() -> int:
    return int(L.load().vnec_device_count_ext())


def _p(a: np.ndarray):
    """Pointer to a's data that keeps a alive for the call (a may be a
    temporary copy)."""
    return a.ctypes.data_as(C.c_void_p)


def _spb_type(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return SpB_FP64
    if dt == np.float32:
        return SpB_FP32
    raise ValueError(f"the VNEC path takes float64 or float32 values, not {dt}")


class Matrix:
    """SpB_Matrix (CSR, FP64 or FP32 values)."""

    def __init__(self, nrows: int, ncols: int, dtype=np.float64):
        lib = init()
        self.h = C.c_void_p()
        self.dtype = np.dtype(dtype)
        L.check(lib.SpB_Matrix_new(C.byref(self.h), _spb_type(dtype), nrows, ncols),
                "SpB_Matrix_new")
        self.nrows, self.ncols = nrows, ncols
        self.nnz = 0

    @property
    def fp32(self) -> bool:
        return self.dtype == np.float32

    @classmethod
    def from_csr(cls, nrows, ncols, ptr, col, val, dtype=np.float64) -> "Matrix":
        m = cls(nrows, ncols, dtype)
        ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=m.dtype)
        f = "SpB_Matrix_import_csr_fp32_ext" if m.fp32 else "SpB_Matrix_import_csr_ext"
        L.check(getattr(L.load(), f)(m.h, _p(ptr), _p(col), _p(val)), f)
        m.nnz = int(ptr[-1])
        return m

    @classmethod
    def from_batch(cls, row_off, col_off, ptr, col_local, val) -> "Matrix":
        """Block-diagonal batch (vnec_batch_import_ext)."""
        row_off = np.ascontiguousarray(row_off, dtype=np.int64)
        col_off = np.ascontiguousarray(col_off, dtype=np.int64)
        m = cls(int(row_off[-1]), int(col_off[-1]))
        ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        col = np.ascontiguousarray(col_local, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        L.check(L.load().vnec_batch_import_ext(m.h, len(row_off) - 1, _p(row_off), _p(col_off),
                                               _p(ptr), _p(col), _p(val)), "vnec_batch_import_ext")
        m.nnz = int(ptr[-1])
        return m

    @classmethod
    def from_coo(cls, nrows, ncols, rows, cols, vals, dtype=np.float64) -> "Matrix":
        m = cls(nrows, ncols, dtype)
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        c = np.ascontiguousarray(cols, dtype=np.uint64)
        v = np.ascontiguousarray(vals, dtype=m.dtype)
        f = "SpB_Matrix_build_FP32" if m.fp32 else "SpB_Matrix_build_FP64"
        L.check(getattr(L.load(), f)(m.h, _p(r), _p(c), _p(v), len(v), 0), f)
        m.nnz = len(v)
        return m

    def free(self):
        if self.h:
            L.check(L.load().SpB_Matrix_free(C.byref(self.h)), "SpB_Matrix_free")
            self.h = C.c_void_p()

    def __del__(self):
        try:
            if self.h:
                L.load().SpB_Matrix_free(C.byref(self.h))
        except Exception:
            pass


class Vector:
    """SpB_Vector (dense, FP64 or FP32 values)."""

    def __init__(self, n: int, dtype=np.float64):
        lib = init()
        self.h = C.c_void_p()
        self.dtype = np.dtype(dtype)
        L.check(lib.SpB_Vector_new(C.byref(self.h), _spb_type(dtype), n), "SpB_Vector_new")
        self.n = n

    @property
    def fp32(self) -> bool:
        return self.dtype == np.float32

    @classmethod
    def from_array(cls, a, dtype=np.float64) -> "Vector":
        a = np.ascontiguousarray(a, dtype=dtype)
        v = cls(len(a), dtype)
        f = "SpB_Vector_import_fp32_ext" if v.fp32 else "SpB_Vector_import_ext"
        L.check(getattr(L.load(), f)(v.h, _p(a)), f)
        return v

    def to_array(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        f = "SpB_Vector_export_fp32_ext" if self.fp32 else "SpB_Vector_export_ext"
        L.check(getattr(L.load(), f)(self.h, _p(out)), f)
        return out

    def set(self, i: int, x: float):
        f = "SpB_Vector_setElement_FP32" if self.fp32 else "SpB_Vector_setElement_FP64"
        L.check(getattr(L.load(), f)(self.h, float(x), i), f)

    def get(self, i: int) -> float:
        out = C.c_float() if self.fp32 else C.c_double()
        f = "SpB_Vector_extractElement_FP32" if self.fp32 else "SpB_Vector_extractElement_FP64"
        L.check(getattr(L.load(), f)(C.byref(out), self.h, i), f)
        return out.value

    def resident(self, on: bool = True):
        L.check(L.load().SpB_Vector_device_resident_ext(self.h, on), "device_resident")

    def device_ptr(self) -> int:
        p = L.load().SpB_Vector_device_ptr_ext(self.h)
        if not p:
            raise RuntimeError(L.last_error())
        return p

    def free(self):
        if self.h:
            L.check(L.load().SpB_Vector_free(C.byref(self.h)), "SpB_Vector_free")
            self.h = C.c_void_p()

    def __del__(self):
        try:
            if self.h:
                L.load().SpB_Vector_free(C.byref(self.h))
        except Exception:
            pass


class DeviceArray:
    """A raw device buffer owned by the library's allocator."""

    def __init__(self, nbytes: int):
        init()
        self.nbytes = nbytes
        self.ptr = L.load().vnec_dev_alloc_ext(nbytes)
        if not self.ptr:
            raise RuntimeError(L.last_error())

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        L.check(L.load().vnec_memcpy_h2d_ext(d.ptr, _p(a), a.nbytes), "h2d")
        return d

    def to_array(self, dtype, n) -> np.ndarray:
        out = np.empty(n, dtype=dtype)
        L.check(L.load().vnec_memcpy_d2h_ext(_p(out), self.ptr, out.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            L.load().vnec_dev_free_ext(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class TuningInfo:
    variant: str
    ipt: int
    lanes: int
    encoding: str
    nrhs: int
    grid: int
    ntiles: int
    dev_bytes: int


# vnec_tuning.lanes flags of L and C plans (include/vnec_ext.h)
LANES_RELABEL = 1  # columns relabelled hot-first, x permuted before each SpMV
LANES_WIDE = 1024  # C only: 512-thread tiles of 512 * ipt items
LANES_GATHER = 2048  # C only, per-tile carries: a thread's x gathers of a tile issued together


class Plan:
    """GPU execution plan (vnec_plan_build_ext).  variant: a variant letter
    or a selection mode ("learned", "autotune", "rule", "reference");
    lanes: the D block width, H's hot entries, or for L and C 64 or the
    flags LANES_RELABEL | LANES_WIDE | LANES_GATHER."""

    def __init__(self, A: Matrix, variant: str | None = "L", ipt: int = 7, lanes: int = 64,
                 encoding: str | None = None, nrhs: int = 1, grid: int = 0):
        init()
        if encoding is None:  # the matrix's own arithmetic
            encoding = "fp32" if A.fp32 else "fp64"
        code = -1 if variant is None else (VNEC_VAR[variant] if variant in VNEC_VAR
                                            else SELECT[variant])
        t = L.Tuning(code, ipt, lanes, VNEC_ENC[encoding], nrhs, grid)
        self.h = C.c_void_p()
        self.A = A
        L.check(L.load().vnec_plan_build_ext(C.byref(self.h), A.h, C.byref(t)),
                "vnec_plan_build_ext")

    def info(self) -> TuningInfo:
        t = L.Tuning()
        nt, db = C.c_int64(), C.c_int64()
        L.check(L.load().vnec_plan_info_ext(self.h, C.byref(t), C.byref(nt), C.byref(db)),
                "vnec_plan_info_ext")
        inv = {v: k for k, v in VNEC_VAR.items()}
        inve = {v: k for k, v in VNEC_ENC.items()}
        return TuningInfo(inv[t.variant], t.ipt, t.lanes, inve[t.encoding], t.nrhs, t.grid,
                          nt.value, db.value)

    def format_bytes(self) -> tuple[int, int]:
        """(main kernel, whole SpMV) bytes the plan's own layout must move
        per SpMV (vnec_plan_bytes_ext)."""
        a, b = C.c_int64(), C.c_int64()
        L.check(L.load().vnec_plan_bytes_ext(self.h, C.byref(a), C.byref(b)), "vnec_plan_bytes_ext")
        return a.value, b.value

    def spmv(self, y: Vector, x: Vector):
        L.check(L.load().vnec_spmv_ext(self.h, y.h, x.h), "vnec_spmv_ext")

    def spmv_dev(self, x_ptr: int, y_ptr: int):
        L.check(L.load().vnec_spmv_dev_ext(self.h, x_ptr, y_ptr), "vnec_spmv_dev_ext")

    def spmv_graph(self, x_ptr: int, y_ptr: int):
        """vnec_spmv_graph_ext: the same SpMV replayed from the plan's
        captured graph for this pointer pair."""
        L.check(L.load().vnec_spmv_graph_ext(self.h, x_ptr, y_ptr), "vnec_spmv_graph_ext")

    def spmm_colmajor_dev(self, x_ptr: int, y_ptr: int):
        """Right-hand sides column-major: nrhs vectors of ncols (x) and of
        nrows (y), one after another."""
        L.check(L.load().vnec_spmm_colmajor_dev_ext(self.h, x_ptr, y_ptr), "vnec_spmm_colmajor_dev_ext")

    def time(self, x_ptr: int, y_ptr: int, iters: int, graph: bool = True,
             flush_bytes: int = 0) -> float:
        ms = C.c_double()
        L.check(L.load().vnec_time_dev_ext(self.h, x_ptr, y_ptr, iters, int(graph), flush_bytes,
                                           C.byref(ms)), "vnec_time_dev_ext")
        return ms.value

    def time_kernels(self, x_ptr: int, y_ptr: int, iters: int) -> tuple[float, float]:
        """(SpMV kernel ms, carry fix-up ms): mean of iters back-to-back launches."""
        a, b = C.c_double(), C.c_double()
        L.check(L.load().vnec_time_kernels_ext(self.h, x_ptr, y_ptr, iters, C.byref(a),
                                               C.byref(b)), "vnec_time_kernels_ext")
        return a.value, b.value

    def free(self):
        if self.h:
            L.check(L.load().vnec_plan_free_ext(C.byref(self.h)), "vnec_plan_free_ext")
            self.h = C.c_void_p()

    def __del__(self):
        try:
            if self.h:
                L.load().vnec_plan_free_ext(C.byref(self.h))
        except Exception:
            pass


# ------------------------------------------------ reference entry points
def ird_vnec(A: Matrix, fp32: bool | None = None) -> int:
    """IRD_VNEC(A, SpB_FP32 or SpB_FP64); the type defaults to A's."""
    fp32 = A.fp32 if fp32 is None else fp32
    return int(L.load().IRD_VNEC(A.h, SpB_FP32 if fp32 else SpB_FP64))


def ird_ext(A: Matrix, P=36, L_=8, thr=0.477) -> tuple[int, float]:
    r, d = C.c_float(), C.c_int()
    L.check(L.load().vnec_ird_ext(A.h, P, L_, thr, C.byref(r), C.byref(d)), "vnec_ird_ext")
    return d.value, r.value


def build_vnec(kind: str, A: Matrix):
    """Build_VNEC_{L, D_FP64, S_FP64} (or the FP32 builders for an FP32 A)."""
    sfx = "FP32" if A.fp32 else "FP64"
    f = {"L": "Build_VNEC_L", "D": f"Build_VNEC_D_{sfx}", "S": f"Build_VNEC_S_{sfx}"}[kind]
    p = getattr(L.load(), f)(A.h)
    if not p:
        raise RuntimeError(f"{f}: {L.last_error()}")
    return p


def spmv_vnec(kind: str, y: Vector, A: Matrix, x: Vector, plan, fp32: bool | None = None):
    """SpMV_VNEC_{L,D,S}_FP64 (or _FP32: by default A's type); returns the
    library's error text ("" on success), as these calls return void."""
    fp32 = A.fp32 if fp32 is None else fp32
    f = f"SpMV_VNEC_{kind}_{'FP32' if fp32 else 'FP64'}"
    getattr(L.load(), f)(y.h, A.h, x.h, C.cast(plan, C.c_void_p))
    err = L.last_error()
    return err


def reference_plan_info(plan) -> TuningInfo:
    """Tuning of the device plan behind a Build_VNEC_* result."""
    lib = L.load()
    h = lib.VNEC_plan_of_ext(C.cast(plan, C.c_void_p))
    if not h:
        raise RuntimeError("not a plan of this library")
    t = L.Tuning()
    nt, db = C.c_int64(), C.c_int64()
    L.check(lib.vnec_plan_info_ext(h, C.byref(t), C.byref(nt), C.byref(db)), "vnec_plan_info_ext")
    inv = {v: k for k, v in VNEC_VAR.items()}
    inve = {v: k for k, v in VNEC_ENC.items()}
    return TuningInfo(inv[t.variant], t.ipt, t.lanes, inve[t.encoding], t.nrhs, t.grid, nt.value,
                      db.value)


def free_vnec(kind: str, plan, fp32: bool = False):
    sfx = "fp32" if fp32 else "fp64"
    f = {"L": "VNEC_L_free_ext", "D": f"VNEC_D_free_{sfx}_ext", "S": f"VNEC_S_free_{sfx}_ext"}[kind]
    L.check(getattr(L.load(), f)(C.cast(plan, C.c_void_p)), f)


def mxv(y: Vector, A: Matrix, x: Vector, plan=None):
    p = None if plan is None else (plan.h if isinstance(plan, Plan) else C.cast(plan, C.c_void_p))
    L.check(L.load().SpB_mxv(y.h, A.h, x.h, p), "SpB_mxv")


# ------------------------------------------------ builder structures
def partition(A: Matrix, P: int, snap: bool = False) -> np.ndarray:
    out = np.zeros((P + 1, 2), dtype=np.int32)
    L.check(L.load().vnec_partition_ext(A.h, P, int(snap), _p(out)), "vnec_partition_ext")
    return out


def ecr(A: Matrix, coords: np.ndarray, method: int = 0):
    coords = np.ascontiguousarray(coords, dtype=np.int32)
    P = coords.shape[0] - 1
    nnz = A.nnz
    e = np.zeros(max(nnz, 1), np.int32)
    nec = np.zeros(P, np.int32)
    off = np.zeros(P + 1, np.int64)
    win = np.zeros(max(nnz, 1), np.int32)
    L.check(L.load().vnec_ecr_ext(A.h, P, _p(coords), method, _p(e), _p(nec), _p(off), _p(win)),
            "vnec_ecr_ext")
    return e[:nnz], nec, off, win[: int(off[-1])]


def s_order(A: Matrix, P: int, L_: int, cut: int = 32) -> dict:
    """VNEC-S task order and reordered matrix at (P, L, cut) (vnec_s_order_ext)."""
    n, nz = A.nrows, A.nnz
    out = {"tasks": np.zeros(max(n, 1), np.int32), "spvv_len": np.zeros(P, np.int32),
           "row_begin": np.zeros(max(n, 1), np.int32), "row_end": np.zeros(max(n, 1), np.int32),
           "s_val": np.zeros(max(nz, 1), np.float64), "s_col": np.zeros(max(nz, 1), np.int32),
           "task_start": np.zeros(P, np.int32), "task_end": np.zeros(P, np.int32)}
    L.check(L.load().vnec_s_order_ext(A.h, P, L_, cut, *(_p(out[k]) for k in
                                      ("tasks", "spvv_len", "row_begin", "row_end", "s_val",
                                       "s_col", "task_start", "task_end"))), "vnec_s_order_ext")
    for k in ("tasks", "row_begin", "row_end"):
        out[k] = out[k][:n]
    for k in ("s_val", "s_col"):
        out[k] = out[k][:nz]
    return out


def d_layout(A: Matrix, P: int, L_: int, with_values: bool = True):
    G = C.c_int64()
    L.check(L.load().vnec_d_layout_ext(A.h, P, L_, None, None, None, C.byref(G)),
            "vnec_d_layout_ext")
    g = G.value
    vrp = np.zeros(A.nrows + 1, np.int32)
    cs = np.zeros(max(g, 1), np.int32)
    va = np.zeros(max(g * L_, 1), np.float64) if with_values else None
    L.check(L.load().vnec_d_layout_ext(A.h, P, L_, _p(vrp), _p(cs),
                                       _p(va) if with_values else None, C.byref(G)),
            "vnec_d_layout_ext")
    return vrp, cs[:g], (va[: g * L_] if with_values else None)


# ------------------------------------------------ graph files
def read_graph(path: str) -> tuple[int, np.ndarray, np.ndarray]:
    """(n, ptr int64, col int32) of a .mtx / .sg file, read as the reference
    harness's Builder does (vnec_read_graph_ext)."""
    lib = L.load()
    n = C.c_int64()
    pp, pc = C.c_void_p(), C.c_void_p()
    L.check(lib.vnec_read_graph_ext(path.encode(), C.byref(n), C.byref(pp), C.byref(pc)),
            "vnec_read_graph_ext")
    try:
        ptr = np.ctypeslib.as_array(C.cast(pp, C.POINTER(C.c_int64)), (n.value + 1,)).copy()
        nz = int(ptr[-1])
        col = (np.ctypeslib.as_array(C.cast(pc, C.POINTER(C.c_int32)), (nz,)).copy() if nz
               else np.zeros(0, np.int32))
    finally:
        lib.vnec_host_free_ext(pp)
        lib.vnec_host_free_ext(pc)
    return n.value, ptr, col


def pattern(A: Matrix) -> tuple[np.ndarray, int]:
    """Variant G's dominant row pattern (column offsets) and the number of
    rows whose columns are a subset of it (vnec_pattern_ext)."""
    pat = np.zeros(32, np.int32)
    L_, nreg = C.c_int(), C.c_int64()
    L.check(L.load().vnec_pattern_ext(A.h, _p(pat), C.byref(L_), C.byref(nreg)), "vnec_pattern_ext")
    return pat[:L_.value].copy(), nreg.value


def pattern_match(A: Matrix, pat) -> tuple[np.ndarray, np.ndarray]:
    """(masks, flags) per row: the pattern positions the row holds, and 1
    where the row is not a subset of the pattern (GPU)."""
    pat = np.ascontiguousarray(pat, np.int32)
    masks = np.zeros(A.nrows, np.uint32)
    flags = np.zeros(A.nrows, np.int32)
    L.check(L.load().vnec_pattern_match_ext(A.h, _p(pat), len(pat), _p(masks), _p(flags)),
            "vnec_pattern_match_ext")
    return masks, flags


def matrix_from_file(path: str, values: str = "row", dtype=np.float64) -> Matrix:
    """SpB_Matrix_read_ext: a_ij = i (\"row\", the harness's values) or 1 (\"one\")."""
    init()
    m = Matrix.__new__(Matrix)
    m.h = C.c_void_p()
    m.dtype = np.dtype(dtype)
    L.check(L.load().SpB_Matrix_read_ext(C.byref(m.h), path.encode(), _spb_type(dtype),
                                         {"row": 0, "one": 1}[values]), "SpB_Matrix_read_ext")
    n, ptr, _ = read_graph(path)
    m.nrows = m.ncols = n
    m.nnz = int(ptr[-1])
    return m


# ------------------------------------------------ autotuner and selector
def candidates() -> list[TuningInfo]:
    """The candidate configurations (vnec_candidates_ext)."""
    arr = (L.Tuning * L.NCAND)()
    n = L.load().vnec_candidates_ext(arr, L.NCAND)
    inv = {v: k for k, v in VNEC_VAR.items()}
    return [TuningInfo(inv[t.variant], t.ipt, t.lanes, "fp64", t.nrhs, t.grid, 0, 0)
            for t in arr[:n]]


def feature_ints(A: Matrix) -> np.ndarray:
    """The 32 integer statistics computed on the GPU (vnec_feature_ints_ext)."""
    init()
    out = np.zeros(L.NFEAT_INTS, np.uint64)
    L.check(L.load().vnec_feature_ints_ext(A.h, _p(out)), "vnec_feature_ints_ext")
    return out


def features_from_ints(ints) -> np.ndarray:
    ints = np.ascontiguousarray(ints, dtype=np.uint64)
    f = np.zeros(L.NFEAT, np.float32)
    L.load().vnec_features_from_ints_ext(_p(ints), _p(f))
    return f


def mask_from_ints(ints) -> int:
    ints = np.ascontiguousarray(ints, dtype=np.uint64)
    return int(L.load().vnec_candidate_mask_from_ints_ext(_p(ints)))


def candidate_mask(A: Matrix) -> int:
    init()
    m = C.c_uint32()
    L.check(L.load().vnec_candidate_mask_ext(A.h, C.byref(m)), "vnec_candidate_mask_ext")
    return int(m.value)


def autotune(A: Matrix, reps: int = 20) -> tuple[np.ndarray, int]:
    """(per-candidate ms, inf where not applicable; index of the fastest)."""
    init()
    t = np.zeros(L.NCAND, np.float64)
    b = C.c_int()
    L.check(L.load().vnec_autotune_ext(A.h, reps, _p(t), C.byref(b)), "vnec_autotune_ext")
    return t, int(b.value)


def select(A: Matrix, mode: str = "learned") -> int:
    init()
    c = C.c_int()
    L.check(L.load().vnec_select_ext(A.h, SELECT[mode], C.byref(c)), "vnec_select_ext")
    return int(c.value)


def _f32(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def mlp_forward(params, X) -> np.ndarray:
    """HIP forward pass: (B, NCAND) outputs of standardised inputs X (B, 32)."""
    init()
    X = _f32(X, (-1, L.NFEAT))
    out = np.zeros((X.shape[0], L.NCAND), np.float32)
    L.check(L.load().vnec_mlp_forward_ext(_p(_f32(params)), _p(X), X.shape[0], _p(out)),
            "vnec_mlp_forward_ext")
    return out


def mlp_grad(params, X, T, M) -> tuple[np.ndarray, float]:
    """HIP backward pass: gradient of sum M (out - T)^2 / sum M, and the loss."""
    init()
    X = _f32(X, (-1, L.NFEAT))
    g = np.zeros(L.MLP_NPARAMS, np.float32)
    loss = C.c_float()
    L.check(L.load().vnec_mlp_grad_ext(_p(_f32(params)), _p(X), _p(_f32(T)), _p(_f32(M)),
                                       X.shape[0], _p(g), C.byref(loss)), "vnec_mlp_grad_ext")
    return g, float(loss.value)


def mlp_train(params, X, T, M, epochs: int, batch: int, lr: float, weight_decay: float = 0.0,
              seed: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """Adam training with the HIP kernels; returns (params, per-epoch loss)."""
    init()
    p = np.array(params, dtype=np.float32, copy=True)
    X = _f32(X, (-1, L.NFEAT))
    hist = np.zeros(max(epochs, 1), np.float32)
    L.check(L.load().vnec_mlp_train_ext(_p(p), _p(X), _p(_f32(T)), _p(_f32(M)), X.shape[0],
                                        epochs, batch, lr, weight_decay, seed, _p(hist)),
            "vnec_mlp_train_ext")
    return p, hist[:epochs]


def mlp_select(params, mu, sd, F, masks) -> np.ndarray:
    """HIP inference kernel: chosen candidate per row of raw features F."""
    init()
    F = _f32(F, (-1, L.NFEAT))
    masks = np.ascontiguousarray(masks, dtype=np.uint32)
    out = np.zeros(F.shape[0], np.int32)
    L.check(L.load().vnec_mlp_select_ext(_p(_f32(params)), _p(_f32(mu)), _p(_f32(sd)), _p(F),
                                         _p(masks), F.shape[0], _p(out)), "vnec_mlp_select_ext")
    return out


def selector_install(params=None, mu=None, sd=None):
    init()
    if params is None:
        L.check(L.load().vnec_selector_install_ext(None, None, None), "vnec_selector_install_ext")
    else:
        L.check(L.load().vnec_selector_install_ext(_p(_f32(params)), _p(_f32(mu)), _p(_f32(sd))),
                "vnec_selector_install_ext")


def selector_info() -> tuple[bool, int]:
    t, n = C.c_int(), C.c_int64()
    L.check(L.load().vnec_selector_info_ext(C.byref(t), C.byref(n)), "vnec_selector_info_ext")
    return bool(t.value), int(n.value)
