"""The learned variant selector: seeded training matrices, the autotuner's
labels, GPU training with the library's own kernels, and evaluation.

    python -m vnec_amd.selector.gen_dataset --start 0 --count 550 --out ds0.npz
    python -m vnec_amd.selector.train --data vnec_amd/selector/dataset.npz
"""
