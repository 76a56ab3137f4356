"""Writes tests/golden/ird_large.json: the reference's IRD on matrices of
more than 2^24 nonzeros, where its FP32 ratio (float)nnz / (float)(G * L)
(src/operation/src/ird_calc.c:103-104) no longer holds every operand
exactly.  The matrices are regenerated from their seeds by vnec_amd.matgen,
so the fixture holds only the generator call and what the reference
computed:

  nrows, nnz      — of the generated CSR (a check that it regenerated alike)
  groups          — v_row_ptr[nrows] of the reference's Build_VNEC_D_FP64
                    (the D grouping IRD counts, P = 36, L = 8)
  ird_printed     — the ratio as IRD_VNEC prints it ("%f", ird_calc.c:105)
  decision        — IRD_VNEC(A, SpB_FP64): 0 = SpB_VNEC_D, 1 = SpB_VNEC_S

Run from the repository root, where the reference library is built
(make -C oracle ref):  python tests/golden/make_ird_golden.py
"""
from __future__ import annotations

import json
import os
import re
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from vnec_amd import matgen as g  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ird_large.json")

CASES = [
    {"name": "stencil27_128", "gen": "stencil27", "args": [128], "seed": 1},
    {"name": "rmat21", "gen": "rmat", "args": [21, 16, 0.57, 0.19, 0.19], "seed": 1},
    {"name": "rmat21_a045", "gen": "rmat", "args": [21, 16, 0.45, 0.22, 0.22], "seed": 7},
    {"name": "banded2M_hb8", "gen": "banded", "args": [2_097_152, 8], "seed": 3},
]


def captured_ird(R) -> tuple[int, str]:
    """IRD_VNEC's decision and the line it prints on fd 1."""
    import ctypes as C

    libc = C.CDLL(None)
    with tempfile.TemporaryFile(mode="w+b") as f:
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(f.fileno(), 1)
        try:
            d = R.ird()
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        f.seek(0)
        text = f.read().decode()
    m = re.search(r"IRD_mat = ([0-9.]+)", text)
    return d, m.group(1)


def main():
    if not O.ref_available():
        raise SystemExit("build the reference first: make -C oracle ref")
    out = []
    for case in CASES:
        m = getattr(g, case["gen"])(*case["args"], seed=case["seed"])
        print(f"{case['name']}: nrows {m.nrows} nnz {m.nnz}", file=sys.stderr, flush=True)
        R = O.RefMatrix(m.nrows, m.ncols, m.ptr, m.col, m.val)
        decision, printed = captured_ird(R)
        r = O.ref()
        p = r.ref_build_d(R.h)
        groups = int(r.ref_d_groups(p, R.h))
        r.ref_free_d(p)
        R.close()
        out.append(dict(case, nrows=int(m.nrows), nnz=int(m.nnz), groups=groups,
                        ird_printed=printed, decision=int(decision)))
        print(f"  groups {groups} IRD {printed} decision {decision}", file=sys.stderr, flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
