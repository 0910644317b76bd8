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


def s_fields_agree(a, b, tol=5e-4, max_flip=2e-3):
    """Two scalar-operator fields S agree: within tol everywhere except at cells whose |grad u|
    lies within rounding of the 0.01*G threshold, where S jumps between 1 and S(n); those may
    differ on at most a fraction max_flip of the cells."""
    d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    far = d > tol
    return far.mean() <= max_flip and (not (~far).any() or d[~far].max() <= tol)


# host restatement of the predicted-threshold band (gmax_mode 2), shared with bench.py
from qp3d.threshold import band, predict_G, s_threshold, steps_in_band  # noqa: E402,F401
