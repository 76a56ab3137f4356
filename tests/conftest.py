import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def golden_names(fp32: bool = False):
    """FP64 fixtures `<name>.npz`, or the FP32 ones `<name>_f32.npz`."""
    names = sorted(os.path.splitext(os.path.basename(p))[0]
                   for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
    return [n for n in names if n.endswith("_f32") == fp32]


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(params=golden_names())
def golden(request):
    d = load_golden(request.param)
    d["name"] = request.param
    return d


@pytest.fixture(params=golden_names(fp32=True))
def golden32(request):
    """Fixtures of the reference's FP32 entry points (values and x FP32,
    L = 16, IRD threshold 0.437)."""
    d = load_golden(request.param)
    d["name"] = request.param
    return d
