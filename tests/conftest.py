import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference")

# On a GPU machine, torch's HIP runtime starts before libofx.so is loaded, so that the tests that
# render or hold device arrays with torch and the library share one runtime (loaded the other way
# round, torch finds no device). Without a GPU this does nothing.
try:
    import torch

    if torch.cuda.is_available():
        torch.cuda.init()
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def load(name):
    """A fixture of the reference as float64 (tests/golden/reference)."""
    return np.load(os.path.join(GOLDEN, name + ".npy")).astype(np.float64)


def meta():
    with open(os.path.join(GOLDEN, "meta.json")) as fh:
        return json.load(fh)


def rms(a, b, m=None):
    if m is None:
        m = np.isfinite(a) & np.isfinite(b)
    return float(np.sqrt(np.mean((a[m] - b[m]) ** 2)))
