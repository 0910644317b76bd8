import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "gpu-c_3dqmodeling_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running check, opt in with QP3D_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("QP3D_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow test; set QP3D_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0)

