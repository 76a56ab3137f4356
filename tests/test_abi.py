"""CPU-side checks of the C-ABI library (no GPU needed): it loads, exports
every function include/*.h declares, and its host object logic matches the
reference's (struct layout prefix, CSR build order, error returns)."""
import ctypes as C
import os
import re
from pathlib import Path

import numpy as np
import pytest

from vnec_amd import _lib as L
from vnec_amd import spb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("OpenSparseBLAS.h", "vnec_ext.h", "vnec_hip.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"typedef[^;]*;", "", txt, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", txt):
            names.add(m.group(1))
    names.discard("if")
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = L.load()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers them all
    assert not [n for n in names if n not in L.SIGNATURES]


class RefMatrixLayout(C.Structure):
    """src/core/include/SpB_Matrix.h:10-17 + SpB_Matrix_CSC_or_CSR.h:12-21."""
    _fields_ = [("element_type", C.c_int), ("format_type", C.c_int), ("row", C.c_uint64),
                ("col", C.c_uint64), ("element_size", C.c_size_t), ("val", C.c_void_p),
                ("indices", C.POINTER(C.c_uint64)), ("ptr", C.POINTER(C.c_uint64)),
                ("ptr_len", C.c_uint64), ("nnz", C.c_uint64), ("capacity", C.c_uint64)]


class RefVectorLayout(C.Structure):
    """src/core/include/SpB_Vector.h:10-15 + SpB_Vector_Dense.h:15-19."""
    _fields_ = [("element_type", C.c_int), ("format_type", C.c_int), ("n", C.c_uint64),
                ("element_size", C.c_size_t), ("values", C.POINTER(C.c_double))]


def _view(h, cls):
    return C.cast(h, C.POINTER(cls)).contents


def test_matrix_build_matches_reference_layout_and_order():
    rng = np.random.default_rng(0)
    n, m, z = 50, 40, 300
    r = rng.integers(0, n, z)
    c = rng.integers(0, m, z)
    v = rng.uniform(-1, 1, z)
    A = spb.Matrix.from_coo(n, m, r, c, v)
    s = _view(A.h, RefMatrixLayout)
    assert (s.element_type, s.format_type, s.row, s.col, s.element_size) == (11, 1, n, m, 8)
    assert s.nnz == z and s.ptr_len == n + 1
    ptr = np.ctypeslib.as_array(s.ptr, shape=(n + 1,))
    ind = np.ctypeslib.as_array(s.indices, shape=(z,))
    val = np.ctypeslib.as_array(C.cast(s.val, C.POINTER(C.c_double)), shape=(z,))
    # sorted by (row, col), duplicates kept in input order (a stable sort)
    order = np.lexsort((np.arange(z), c, r))
    assert np.array_equal(ind, c[order]) and np.array_equal(val, v[order])
    expect_ptr = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=n))])
    assert np.array_equal(ptr, expect_ptr)  # trailing empty rows filled too


def test_vector_layout_and_elements():
    x = spb.Vector(10)
    s = _view(x.h, RefVectorLayout)
    assert (s.element_type, s.format_type, s.n, s.element_size) == (11, 0, 10, 8)
    vals = np.ctypeslib.as_array(s.values, shape=(10,))
    assert np.all(vals == 0)
    x.set(3, 2.5)
    assert x.get(3) == 2.5 and vals[3] == 2.5
    with pytest.raises(RuntimeError):
        x.set(10, 1.0)


def test_error_paths_return_failure_not_exit():
    lib = spb.init()
    h = C.c_void_p()
    assert lib.SpB_Matrix_new(C.byref(h), 4, 3, 3) == 1  # SpB_INT32 is not on the path
    assert "SpB_FP64 or SpB_FP32" in L.last_error()
    assert lib.SpB_Vector_new(C.byref(h), 4, 3) == 1
    assert "SpB_FP64 or SpB_FP32" in L.last_error()
    A = spb.Matrix(4, 4)
    bad_r = np.array([0, 9], np.uint64)
    cc = np.array([0, 1], np.uint64)
    vv = np.array([1.0, 2.0])
    assert lib.SpB_Matrix_build_FP64(A.h, bad_r.ctypes.data, cc.ctypes.data, vv.ctypes.data, 2,
                                     0) == 1
    assert "out of range" in L.last_error()
    assert lib.SpB_Matrix_build_FP64(A.h, cc.ctypes.data, cc.ctypes.data, vv.ctypes.data, 2,
                                     5) == 1
    x, y = spb.Vector(5), spb.Vector(4)
    assert lib.SpB_mxv(y.h, A.h, x.h, None) == 1  # dimension mismatch
    assert "dimension" in L.last_error()


def test_plan_without_device_fails_loudly():
    if spb.device_count() > 0:
        pytest.skip("a GPU is present")
    A = spb.Matrix.from_csr(2, 2, [0, 1, 2], [0, 1], [1.0, 2.0])
    with pytest.raises(RuntimeError, match="no HIP device"):
        spb.Plan(A)


@pytest.mark.parametrize("ptr,col,msg", [
    ([0, 1, 2], [0, 5], "out of range"),
    ([0, 2, 1, 3], [0, 1, 2], "monotone"),
])
def test_invalid_structure_is_refused_at_import(ptr, col, msg):
    n = len(ptr) - 1
    with pytest.raises(RuntimeError, match=msg):
        spb.Matrix.from_csr(n, 3, np.array(ptr, np.int64), np.array(col, np.int32),
                            np.ones(len(col)))


@pytest.mark.parametrize("which,msg", [("col", "not below ncols"), ("ptr", "decrease")])
def test_invalid_structure_is_refused_before_any_kernel(which, msg):
    """Callers that fill the public CSR fields themselves (as the reference
    harness downcasts handles) bypass import: the device mirror re-checks
    every column and row pointer on the host before any kernel reads them."""
    A = spb.Matrix.from_csr(3, 3, np.array([0, 1, 2, 3], np.int64), np.array([0, 1, 2], np.int32),
                            np.ones(3))
    s = _view(A.h, RefMatrixLayout)
    if which == "col":
        s.indices[1] = 7
    else:
        s.ptr[1] = 2
        s.ptr[2] = 1
    with pytest.raises(RuntimeError, match=msg):
        spb.partition(A, 2)


def test_fp32_objects_and_type_checks_on_the_host():
    """FP32 matrices and vectors keep the reference layouts (element_type
    SpB_FP32 = 10, element_size 4, float values); mixing value types is
    refused before any device work."""
    import numpy as np_
    rng = np_.random.default_rng(1)
    n, m, z = 30, 20, 150
    r, c = rng.integers(0, n, z), rng.integers(0, m, z)
    v = rng.uniform(-1, 1, z).astype(np_.float32)
    A = spb.Matrix.from_coo(n, m, r, c, v, dtype=np_.float32)
    s = _view(A.h, RefMatrixLayout)
    assert (s.element_type, s.format_type, s.element_size) == (10, 1, 4)
    order = np_.lexsort((np_.arange(z), c, r))
    vals = np_.ctypeslib.as_array(C.cast(s.val, C.POINTER(C.c_float)), shape=(z,))
    assert np_.array_equal(vals, v[order])
    x = spb.Vector(m, np_.float32)
    sv = _view(x.h, RefVectorLayout)
    assert (sv.element_type, sv.element_size) == (10, 4)
    x.set(3, 1.5)
    assert x.get(3) == 1.5
    assert np_.ctypeslib.as_array(C.cast(sv.values, C.POINTER(C.c_float)), shape=(m,))[3] == 1.5
    lib = L.load()
    assert lib.SpB_Vector_setElement_FP64(x.h, 1.0, 0) == 1 and "FP32 values" in L.last_error()
    out = C.c_double()
    assert lib.SpB_Vector_extractElement_FP64(C.byref(out), x.h, 0) == 1
    A64 = spb.Matrix.from_coo(n, m, r, c, v.astype(np_.float64))
    assert lib.SpB_Matrix_build_FP32(A64.h, None, None, None, 0, 0) == 1
    assert "another value type" in L.last_error()
    assert not lib.Build_VNEC_D_FP32(A64.h) and "FP64 values" in L.last_error()
    assert not lib.Build_VNEC_S_FP32(A64.h) and "FP64 values" in L.last_error()
    assert not lib.Build_VNEC_D_FP64(A.h) and "FP32 values" in L.last_error()
    assert not lib.Build_VNEC_S_FP64(A.h) and "FP32 values" in L.last_error()
    y64 = spb.Vector(n)
    assert lib.SpB_mxv(y64.h, A.h, x.h, None) == 1 and "FP32" in L.last_error()
    assert lib.IRD_VNEC(A.h, 4) == 1 and "SpB_FP64 or SpB_FP32" in L.last_error()


def test_vector_extract_tuples():
    """SpB_Vector_extractTuples_FP64/FP32 on dense vectors (reference
    SpB_Vector_Dense.c:162-184): *p_nvals must equal n; (i, v[i]) for all i."""
    import numpy as np_
    lib = L.load()
    n = 7
    x = spb.Vector(n)
    for i in range(n):
        x.set(i, 0.5 * i - 1.0)
    I = (C.c_uint64 * n)()
    X = (C.c_double * n)()
    nv = C.c_uint64(n)
    assert lib.SpB_Vector_extractTuples_FP64(I, X, C.byref(nv), x.h) == 0
    assert list(I) == list(range(n)) and list(X) == [0.5 * i - 1.0 for i in range(n)]
    X2 = (C.c_double * n)()
    assert lib.SpB_Vector_extractTuples_FP64(None, X2, C.byref(nv), x.h) == 0
    assert list(X2) == list(X)
    nv_bad = C.c_uint64(n - 1)
    assert lib.SpB_Vector_extractTuples_FP64(I, X, C.byref(nv_bad), x.h) == 1
    assert "p_nvals" in L.last_error()
    X32 = (C.c_float * n)()
    assert lib.SpB_Vector_extractTuples_FP32(I, X32, C.byref(nv), x.h) == 1  # FP64 vector
    y = spb.Vector(n, np_.float32)
    y.set(2, 2.5)
    assert lib.SpB_Vector_extractTuples_FP32(None, X32, C.byref(nv), y.h) == 0
    assert list(X32) == [0, 0, 2.5, 0, 0, 0, 0]


def test_cpp_overloads_compile(tmp_path):
    """The header's C++ overloads resolve by value type (g++, syntax only)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "ov.cpp"
    src.write_text(
        '#include "OpenSparseBLAS.h"\n'
        "SpB_Info f(SpB_Vector v, SpB_Vector w) {\n"
        "  double d = 0; float s = 0; SpB_Index I[4], n = 4; double X[4]; float Y[4];\n"
        "  SpB_Info r = SpB_Vector_setElement(v, 1.0, 0);\n"
        "  r = (SpB_Info)(r | SpB_Vector_setElement(w, 1.0f, 0));\n"
        "  r = (SpB_Info)(r | SpB_Vector_extractElement(&d, v, 0));\n"
        "  r = (SpB_Info)(r | SpB_Vector_extractElement(&s, w, 0));\n"
        "  r = (SpB_Info)(r | SpB_Vector_extractTuples(I, X, &n, v));\n"
        "  return (SpB_Info)(r | SpB_Vector_extractTuples(I, Y, &n, w));\n"
        "}\n")
    inc = Path(__file__).resolve().parents[1] / "include"
    subprocess.run([gxx, "-fsyntax-only", "-Wall", "-Werror", f"-I{inc}", str(src)], check=True)


def test_column_sort_argument_checks():
    """The variant C builders reject shapes they cannot store before any
    device work: packed absolute columns wider than the shift, a packed
    shift beyond 31, and span queries over no tiles."""
    lib = L.load()
    span = C.c_int32(0)
    # ncols 2^21 does not fit a 20-bit packed column without a per-tile base
    assert lib.vnec_hip_csort(None, 4, None, None, 100, 1 << 21, 20, None, None, None, None, None,
                              None, None) != 0
    assert lib.vnec_hip_csort(None, 4, None, None, 100, 1000, 32, None, None, None, None, None,
                              None, None) != 0
    assert lib.vnec_hip_tile_colspan(None, 0, None, None, C.byref(span), None) != 0
    assert lib.vnec_hip_tile_colspan(None, 4, None, None, None, None) != 0


# include/OpenSparseBLAS.h:78-139 in the VNEC repository: the enumerators in
# order, numbered from 0 (no explicit values there).
_BINARYOPS = (["SpB_BINARYOP_NULL", "SpB_BINARYOP_BOOL_ST", "SpB_LOR", "SpB_LAND", "SpB_LXOR",
               "SpB_LXNOR", "SpB_BINARYOP_BOOL_ED"]
              + [n for t in ("INT8", "INT16", "INT32", "INT64", "UINT8", "UINT16", "UINT32",
                             "UINT64")
                 for n in (f"SpB_BINARYOP_{t}_ST", f"SpB_PLUS_{t}", f"SpB_MINUS_{t}",
                           f"SpB_DIV_{t}", f"SpB_BINARYOP_{t}_ED")]
              + ["SpB_BINARYOP_FP32_ST", "SpB_PLUS_FP32", "SpB_TIMES_FP32", "SpB_DIV_FP32",
                 "SpB_MAX_FP32", "SpB_MINUS_FP32", "SpB_BINARYOP_FP32_ED",
                 "SpB_BINARYOP_FP64_ST", "SpB_PLUS_FP64", "SpB_MINUS_FP64", "SpB_DIV_FP64",
                 "SpB_BINARYOP_FP64_ED"])


def _enum_values(header_dir, tmp_path, tag):
    import subprocess
    src = tmp_path / f"enum_{tag}.c"
    src.write_text("#include <stdio.h>\n#include \"OpenSparseBLAS.h\"\nint main(void){\n"
                   + "".join(f'printf("%d\\n", (int){n});\n' for n in _BINARYOPS)
                   + "return 0;}\n")
    exe = tmp_path / f"enum_{tag}"
    subprocess.run(["gcc", "-std=gnu11", "-w", "-I", str(header_dir), "-o", str(exe), str(src)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return [int(v) for v in out.split()]


def test_binaryop_enum_matches_the_reference(tmp_path):
    """Every SpB_BinaryOp enumerator compiles against the drop-in header with
    the reference's value; where the reference's header is present, the two
    headers give the same values."""
    ours = _enum_values(os.path.join(ROOT, "include"), tmp_path, "ours")
    assert ours == list(range(len(_BINARYOPS))) and len(ours) == 59
    ref_inc = "/root/reference/include"
    if os.path.exists(os.path.join(ref_inc, "OpenSparseBLAS.h")):
        assert _enum_values(ref_inc, tmp_path, "ref") == ours


def test_build_refuses_every_operator_but_null():
    """SpB_Matrix_build_* keeps duplicates with SpB_BINARYOP_NULL and fails
    for any other operator (SpB_Matrix_CSC_or_CSR.c:197,252)."""
    lib = spb.init()
    rr = np.array([0, 0], np.uint64)
    cc = np.array([1, 1], np.uint64)
    vv = np.array([1.0, 2.0])
    for op in (1, 55, 58):  # SpB_BINARYOP_BOOL_ST, SpB_PLUS_FP64, SpB_BINARYOP_FP64_ED
        A = spb.Matrix(2, 2)
        assert lib.SpB_Matrix_build_FP64(A.h, rr.ctypes.data, cc.ctypes.data, vv.ctypes.data, 2,
                                         op) == 1
        assert "SpB_BINARYOP_NULL" in L.last_error()
        A.free()


def test_noalloc_spmv_checks_sizes_before_any_device_work():
    """vnec_csr_spmv_noalloc_ext refuses what its fixed scratch cannot hold
    (and malformed sizes) with a message, before touching a device."""
    lib = L.load()
    # 65536 tiles of 1792 merge items: nrows + nnz <= 117,440,512
    assert lib.vnec_csr_spmv_noalloc_ext(60_000_000, 60_000_000, 57_440_513, None, None, None, None,
                                         None, None) == 1
    assert "scratch holds 65536" in L.last_error()
    assert lib.vnec_csr_spmv_noalloc_ext(-1, 4, 0, None, None, None, None, None, None) == 1
    assert "negative" in L.last_error()
    assert lib.vnec_csr_spmv_noalloc_ext(0, 4, 3, None, None, None, None, None, None) == 1
    assert "no rows" in L.last_error()
    assert lib.vnec_csr_spmv_noalloc_ext(10, 10, 20, None, None, None, None, None, None) == 1
    assert "NULL array" in L.last_error()
    # an empty matrix is a no-op
    assert lib.vnec_csr_spmv_noalloc_ext(0, 0, 0, None, None, None, None, None, None) == 0


def test_python_constants_match_the_headers():
    """The Python mirror's constants (lanes flags, candidate and feature
    counts, variants, encodings, the wide block) equal the C headers' own."""
    txt = ""
    for h in ("vnec_ext.h", "vnec_hip.h"):
        txt += open(os.path.join(ROOT, "include", h)).read()

    def define(name):
        m = re.search(r"#define\s+" + name + r"\s+(\d+)", txt)
        assert m, name
        return int(m.group(1))

    def enum(name):
        m = re.search(r"\b" + name + r"\s*=\s*(-?\d+)", txt)
        assert m, name
        return int(m.group(1))

    assert spb.LANES_RELABEL == define("VNEC_LANES_RELABEL")
    assert spb.LANES_WIDE == define("VNEC_LANES_WIDE")
    assert spb.LANES_GATHER == define("VNEC_LANES_GATHER")
    assert L.NCAND == define("VNEC_NCAND")
    assert L.NFEAT_INTS == define("VNEC_NFEAT_INTS")
    for letter, name in (("L", "VNEC_VAR_L"), ("E", "VNEC_VAR_E"), ("S", "VNEC_VAR_S"), ("D", "VNEC_VAR_D"),
                         ("H", "VNEC_VAR_H"), ("G", "VNEC_VAR_G"), ("C", "VNEC_VAR_C")):
        assert spb.VNEC_VAR[letter] == enum(name), letter
    for enc, name in (("fp64", "VNEC_ENC_FP64"), ("hilo", "VNEC_ENC_HILO"), ("hi", "VNEC_ENC_HI"),
                      ("fp32", "VNEC_ENC_FP32")):
        assert spb.VNEC_ENC[enc] == enum(name), enc
    for mode, name in (("learned", "VNEC_SELECT_LEARNED"), ("autotune", "VNEC_SELECT_AUTOTUNE"),
                       ("rule", "VNEC_SELECT_RULE"), ("reference", "VNEC_SELECT_REFERENCE")):
        assert spb.SELECT[mode] == enum(name), mode
