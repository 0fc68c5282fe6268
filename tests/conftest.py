import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libannf_hip.so")


def golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


TREE_FIXTURES = ["tree_gauss.npz", "tree_ties.npz", "tree_dups.npz", "tree_1d.npz"]
ROW_FIXTURES = ["rows_k1.npz", "rows_k3.npz", "rows_k5_noprop.npz", "rows_k4_nosel.npz", "rows_k2_shift.npz",
                "rows_cfg0_shape.npz"]


@pytest.fixture(scope="session")
def gpu_available():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return True
