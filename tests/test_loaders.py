"""The library's graph loaders (vnec_read_graph_ext, C) against the CSR the
reference harness's own Builder builds from the same files (fixtures in
tests/golden/graphs/, written by tests/golden/make_loader_golden.py from
oracle/_ref/ref_loader), and their error paths.  CPU only."""
import os

import numpy as np
import pytest

from vnec_amd import spb

GRAPHS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "graphs")


def _files():
    return sorted(f for f in os.listdir(GRAPHS) if f.endswith((".mtx", ".sg")))


@pytest.fixture(scope="module")
def expected():
    with np.load(os.path.join(GRAPHS, "loaders.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", _files())
def test_loader_matches_reference_builder(name, expected):
    n, ptr, col = spb.read_graph(os.path.join(GRAPHS, name))
    key = name.replace(".", "_")
    assert n == int(expected[key + "__n"])
    np.testing.assert_array_equal(ptr, expected[key + "__ptr"])
    np.testing.assert_array_equal(col, expected[key + "__col"])


def test_mtx_rules_by_hand(tmp_path):
    """1-based ids, a symmetric file adds reversed edges, repeats and
    self-loops dropped, rows sorted."""
    p = tmp_path / "t.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% c\n4 4 5\n"
                 "2 1\n3 1\n3 1\n4 4\n4 2\n")
    n, ptr, col = spb.read_graph(str(p))
    assert n == 4
    rows = [sorted(col[ptr[i]:ptr[i + 1]].tolist()) for i in range(n)]
    assert rows == [[1, 2], [0, 3], [0], [1]]


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix array real general\n2 2\n", "coordinate"),
    ("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1 0\n", "field"),
    ("%%MatrixMarket matrix coordinate real hermitian\n2 2 1\n1 1 1\n", "symmetry"),
    ("%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1\n", "square"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 3 1\n", "out of range"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1\n2 2 1\n", "more entries"),
    ("MatrixMarket matrix coordinate real general\n2 2 0\n", "MatrixMarket"),
])
def test_mtx_errors_return_failure(tmp_path, text, msg):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(RuntimeError, match=msg):
        spb.read_graph(str(p))


def test_sg_errors_return_failure(tmp_path):
    p = tmp_path / "bad.sg"
    p.write_bytes(b"\x00" + (3).to_bytes(8, "little") + (2).to_bytes(8, "little")
                  + b"".join(v.to_bytes(8, "little") for v in (0, 2, 3)) + (5).to_bytes(8, "little"))
    with pytest.raises(RuntimeError, match="truncated|out of range"):
        spb.read_graph(str(p))
    with pytest.raises(RuntimeError, match="suffix"):
        spb.read_graph(str(tmp_path / "graph.txt"))
    with pytest.raises(RuntimeError, match="couldn't open"):
        spb.read_graph(str(tmp_path / "missing.mtx"))


def test_matrix_from_file_has_the_harness_values():
    A = spb.matrix_from_file(os.path.join(GRAPHS, "pdf_test.mtx"), "row")
    n, ptr, col = spb.read_graph(os.path.join(GRAPHS, "pdf_test.mtx"))
    assert (A.nrows, A.nnz) == (n, int(ptr[-1]))
