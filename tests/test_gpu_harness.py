"""The C port of the reference's VNEC test flow (tests/harness/vnec_harness.c)
on the GPU: known-answer input a_ij = i, x_j = j on graph files read as the
reference harness reads them, the harness's dispatch (nrows < 3e6 -> IRD_VNEC
-> D or S, else L), 1 warm-up + 40 timed SpMVs, and its per-row relative
check (1e-8 FP64, 1e-4 FP32) on every row.

Inputs: the reference's two .mtx fixtures and the loader fixtures
(tests/golden/graphs), plus .sg files written here in the on-disk format
of utilities/writer.hpp:12-49: an R-MAT graph, a 27-point stencil and a
3,000,001-row banded ring that takes the harness's L branch.  The dispatch
the harness reports is checked against the oracle's IRD decision.
"""
import glob
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
from vnec_amd import matgen as g
from vnec_amd import spb

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "harness", "vnec_harness")
GRAPHS = os.path.join(ROOT, "tests", "golden", "graphs")
RESULT = re.compile(r"RESULT (\S+) (fp64|fp32) ([LDS]) nrows=(\d+) nnz=(\d+) ms=([\d.]+) "
                    r"bad=(\d+) max_rel=(\S+)")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if spb.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    if not os.path.exists(HARNESS):
        pytest.fail(f"{HARNESS} missing: run `make -C tests/harness`")


def write_sg(path, n, ptr, col):
    """Undirected-layout .sg: bool directed, u64 edges, u64 nodes, u64
    offsets[n + 1], u64 destinations[edges]."""
    with open(path, "wb") as f:
        f.write(struct.pack("<?QQ", False, int(ptr[-1]), n))
        f.write(np.asarray(ptr, "<u8").tobytes())
        f.write(np.asarray(col, "<u8").tobytes())


def run(paths, fp32=False, resident=False, timeout=300):
    cmd = [HARNESS] + (["--fp32"] if fp32 else []) + (["--resident"] if resident else []) + list(paths)
    print("harness:", " ".join(os.path.basename(c) for c in cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    rows = [RESULT.search(line) for line in p.stdout.splitlines()]
    rows = [m.groups() for m in rows if m]
    return p.returncode, rows, p.stdout + p.stderr


def ird_letter(ptr, col, fp32):
    dec = O.ird(ptr, col, 36, 16 if fp32 else 8, 0.437 if fp32 else 0.477)[0]
    return "D" if dec == 0 else "S"


FILES = sorted(glob.glob(os.path.join(GRAPHS, "*.mtx")) + glob.glob(os.path.join(GRAPHS, "*.sg")))


@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_harness_on_fixture_graphs(fp32):
    rc, rows, out = run(FILES, fp32)
    assert rc == 0, out[-3000:]
    assert len(rows) == len(FILES)
    for path, typ, var, nrows, nnz, ms, bad, _ in rows:
        assert typ == ("fp32" if fp32 else "fp64")
        assert int(bad) == 0, (path, out[-2000:])
        n, ptr, col = spb.read_graph(path)
        assert int(nrows) == n and int(nnz) == int(ptr[-1])
        assert var == ird_letter(ptr, col, fp32), path


@pytest.fixture(scope="module")
def generated(tmp_path_factory):
    d = tmp_path_factory.mktemp("sg")
    out = {}
    m = g.rmat(15, 16, seed=41)
    out["rmat15"] = str(d / "rmat15.sg")
    write_sg(out["rmat15"], m.nrows, m.ptr, m.col)
    m = g.stencil27(40, seed=42)
    out["stencil40"] = str(d / "stencil40.sg")
    write_sg(out["stencil40"], m.nrows, m.ptr, m.col)
    n = 3_000_001  # banded ring, 4 neighbours per row: the harness's L branch
    offs = np.array([-2, -1, 1, 2])
    col = np.sort((np.arange(n)[:, None] + offs[None, :]) % n, axis=1).ravel()
    ptr = np.arange(0, 4 * n + 1, 4, dtype=np.int64)
    out["ring3m"] = str(d / "ring3m.sg")
    write_sg(out["ring3m"], n, ptr, col)
    return out


@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_harness_dispatch_and_known_answers(generated, fp32):
    paths = [generated[k] for k in ("rmat15", "stencil40", "ring3m")]
    rc, rows, out = run(paths, fp32)
    assert rc == 0, out[-3000:]
    assert len(rows) == 3
    got = {os.path.basename(r[0]): r for r in rows}
    assert got["ring3m.sg"][2] == "L"
    for k in ("rmat15.sg", "stencil40.sg"):
        n, ptr, col = spb.read_graph(os.path.join(os.path.dirname(paths[0]), k))
        assert got[k][2] == ird_letter(ptr, col, fp32), k
    for r in rows:
        assert int(r[6]) == 0, (r, out[-2000:])
        print("harness timing:", os.path.basename(r[0]), r[1], r[2], "ms per SpMV", r[5])


def test_harness_resident_vectors(generated):
    rc, rows, out = run([generated["rmat15"]], resident=True)
    assert rc == 0 and len(rows) == 1 and int(rows[0][6]) == 0, out[-2000:]
