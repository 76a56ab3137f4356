"""Writes the graph-loader fixtures: input files under tests/golden/graphs/
and, in tests/golden/graphs/loaders.npz, the CSR that the reference harness's own
Builder (utilities/builder.hpp + reader.hpp, compiled by oracle/Makefile
into oracle/_ref/ref_loader) builds from each of them.

Inputs: the reference's two MatrixMarket fixtures
(data/graph/pdf_test/pdf_test.mtx, data/graph/grb_ppt/grb_ppt.mtx, copied
as data) and small synthetic files covering what the readers handle:
comments and blank lines, pattern / real / integer fields, general /
symmetric / skew-symmetric files, repeated entries, self-loops, empty rows,
and serialized .sg graphs (undirected, and directed with incoming arrays).

    make -C oracle && python tests/golden/make_loader_golden.py
"""
from __future__ import annotations

import os
import shutil
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GRAPHS = os.path.join(HERE, "graphs")
REF_LOADER = os.path.join(ROOT, "oracle", "_ref", "ref_loader")
REF_DATA = "/root/reference/data/graph"


def write_mtx(path, field, symmetry, n, entries, comments=("% generated",), blank=False):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {symmetry}\n")
        for c in comments:
            f.write(c + "\n")
        f.write(f"{n} {n} {len(entries)}\n")
        for k, (u, v) in enumerate(entries):
            if field == "pattern":
                f.write(f"{u} {v}\n")
            elif field == "integer":
                f.write(f"{u} {v} {k % 5 - 2}\n")
            else:
                f.write(f"{u} {v} {0.25 * k - 1.5}\n")
            if blank and k % 7 == 3:
                f.write("\n")


def write_sg(path, n, ptr, col, directed=False):
    with open(path, "wb") as f:
        f.write(struct.pack("<?QQ", directed, len(col), n))
        f.write(np.asarray(ptr, "<u8").tobytes())
        f.write(np.asarray(col, "<u8").tobytes())
        if directed:  # incoming arrays (transpose)
            order = np.lexsort((np.repeat(np.arange(n), np.diff(ptr)), col))
            rows = np.repeat(np.arange(n), np.diff(ptr))[order]
            inptr = np.concatenate([[0], np.cumsum(np.bincount(col, minlength=n))])
            f.write(np.asarray(inptr, "<u8").tobytes())
            f.write(np.asarray(rows, "<u8").tobytes())


def make_inputs():
    os.makedirs(GRAPHS, exist_ok=True)
    for name in ("pdf_test", "grb_ppt"):
        shutil.copy(os.path.join(REF_DATA, name, name + ".mtx"), os.path.join(GRAPHS, name + ".mtx"))
    rng = np.random.default_rng(7)
    n = 40
    e = [(int(a), int(b)) for a, b in rng.integers(1, n + 1, (200, 2))]
    e += e[:15] + [(3, 3), (9, 9), (n, n)]            # repeats and self-loops
    rng.shuffle(e)
    write_mtx(os.path.join(GRAPHS, "general_real.mtx"), "real", "general", n, e,
              comments=("% a comment", "%another"), blank=True)
    write_mtx(os.path.join(GRAPHS, "symmetric_pattern.mtx"), "pattern", "symmetric", n,
              [(max(a, b), min(a, b)) for a, b in e[:120]])
    write_mtx(os.path.join(GRAPHS, "skew_integer.mtx"), "integer", "skew-symmetric", n, e[:90])
    # rows 0..9 and the last rows empty
    write_mtx(os.path.join(GRAPHS, "empty_rows.mtx"), "pattern", "general", 60,
              [(int(a), int(b)) for a, b in rng.integers(11, 50, (80, 2))])
    m = 30
    ptr = np.concatenate([[0], np.cumsum(rng.integers(0, 6, m))]).astype(np.int64)
    col = np.concatenate([np.sort(rng.choice(m, k, replace=False)) for k in np.diff(ptr)]).astype(np.int64)
    write_sg(os.path.join(GRAPHS, "undirected.sg"), m, ptr, col, directed=False)
    write_sg(os.path.join(GRAPHS, "directed.sg"), m, ptr, col, directed=True)


def reference_csr(path):
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "csr.bin")
        subprocess.run([REF_LOADER, path, out], check=True, capture_output=True)
        raw = np.fromfile(out, dtype="<u8")
    n, m = int(raw[0]), int(raw[1])
    return n, raw[2:3 + n].astype(np.int64), raw[3 + n:3 + n + m].astype(np.int32)


def main():
    make_inputs()
    out = {}
    for f in sorted(x for x in os.listdir(GRAPHS) if x.endswith((".mtx", ".sg"))):
        n, ptr, col = reference_csr(os.path.join(GRAPHS, f))
        key = f.replace(".", "_")
        out[key + "__n"] = np.int64(n)
        out[key + "__ptr"] = ptr
        out[key + "__col"] = col
        print(f, n, int(ptr[-1]))
    np.savez_compressed(os.path.join(GRAPHS, "loaders.npz"), **out)


if __name__ == "__main__":
    main()
