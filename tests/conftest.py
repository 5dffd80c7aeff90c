import gzip
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def golden_dir(tmp_path_factory):
    """Decompressed copies of tests/golden/*.gz."""
    d = tmp_path_factory.mktemp("golden")
    for f in os.listdir(GOLDEN):
        if f.endswith(".gz"):
            with gzip.open(os.path.join(GOLDEN, f), "rb") as g, open(d / f[:-3], "wb") as o:
                shutil.copyfileobj(g, o)
    return d


def rel_l2(a, b):
    """Relative L2 difference of two arrays (per full field)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    return num / den if den > 0 else num


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
