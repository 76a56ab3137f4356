"""The trained selector model as committed files: model.npz (parameters,
feature standardisation, training summary) and the C source compiled into
the library (csrc/host/spb_selector_model.c), written together.
"""
from __future__ import annotations

import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NPZ = os.path.join(HERE, "model.npz")
C_SRC = os.path.join(os.path.dirname(HERE), "csrc", "host", "spb_selector_model.c")


def _literal(v: float) -> str:
    s = f"{v:.9g}"
    return s + (".0f" if s.lstrip("-").isdigit() else "f")


def _floats(a) -> str:
    vals = [_literal(float(v)) for v in np.asarray(a, dtype=np.float32)]
    return ",\n".join("    " + ", ".join(vals[i:i + 8]) for i in range(0, len(vals), 8))


def write(params, mu, sd, n_train: int, summary: dict, npz: str = NPZ, c_src: str = C_SRC):
    params = np.asarray(params, np.float32)
    mu = np.asarray(mu, np.float32)
    sd = np.asarray(sd, np.float32)
    np.savez(npz, params=params, mu=mu, sd=sd, n_train=np.int64(n_train),
             summary=np.array(json.dumps(summary, sort_keys=True)))
    with open(c_src, "w") as f:
        f.write("/*\n * spb_selector_model.c — the selector model compiled into the library.\n"
                " * Written by vnec_amd/selector/train.py (MLP trained on the GPU with the\n"
                " * library's own kernels on the autotuner's labels); do not edit.\n */\n"
                '#include "spb_internal.h"\n\n'
                "const int vnec_model_trained = 1;\n"
                f"const int64_t vnec_model_train_samples = {int(n_train)};\n"
                f"const float vnec_model_params[VNEC_MLP_NPARAMS] = {{\n{_floats(params)}\n}};\n"
                f"const float vnec_model_mu[VNEC_NFEAT] = {{\n{_floats(mu)}\n}};\n"
                f"const float vnec_model_sd[VNEC_NFEAT] = {{\n{_floats(sd)}\n}};\n")


def load(npz: str = NPZ):
    """(params, mu, sd) of the committed model."""
    with np.load(npz, allow_pickle=False) as z:
        return z["params"], z["mu"], z["sd"]


def summary(npz: str = NPZ) -> dict:
    with np.load(npz, allow_pickle=False) as z:
        return json.loads(str(z["summary"]))


def parse_c(c_src: str = C_SRC):
    """(trained, params, mu, sd) as written in the C source."""
    text = open(c_src).read()
    trained = int(re.search(r"vnec_model_trained = (\d+)", text).group(1))

    def arr(name):
        body = re.search(name + r"\[[A-Z_]+\] = \{(.*?)\};", text, re.S).group(1)
        return np.array([float(t.rstrip("f")) for t in re.findall(r"[-+0-9.eE]+f?", body)],
                        dtype=np.float32)
    return trained, arr("vnec_model_params"), arr("vnec_model_mu"), arr("vnec_model_sd")
