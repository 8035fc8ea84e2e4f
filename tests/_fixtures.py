"""Shared helpers of the test-suite: golden fixtures and tolerances."""
import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MISSING = 1.0e15
# FP64 radiance parity bound vs the reference
RTOL_FP64 = 1.0e-9


def golden_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    for k in ("km", "nmom", "npol", "meta_nstreams", "meta_vector", "meta_plane_parallel", "out_rc"):
        d[k] = int(d[k])
    return d


def max_rel(a, b, missing=MISSING):
    """Largest relative difference over non-MISSING elements; MISSING must agree."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    ma = np.abs(a / missing - 1) < 1e-3
    mb = np.abs(b / missing - 1) < 1e-3
    assert (ma == mb).all(), "MISSING pattern differs"
    ok = ~ma
    if not ok.any():
        return 0.0
    return float(np.max(np.abs(a[ok] - b[ok]) / np.maximum(np.abs(b[ok]), 1e-300)))
