"""qp3d: MI355X-native 3D pure-qP VTI forward modeling (Python side of the C ABI).

    import qp3d
    s = qp3d.config_setup(hx=61, hy=61, nz0=61, pml=20, nt=200, seed=0)
    rec = qp3d.modeling3d(s)            # the reference's modeling3d() on the GPU

qp3d.autograd (imported separately, it needs torch) holds the torch.autograd shim.
"""
from ._lib import EXPORTED, LIB_PATH, NTILES, Qp3dError, load
from .plan import Batch, Fwi, Plan, modeling3d
from .setup import (Setup, build_setup, config_setup, fd_tables, file_to_grid, genmodel, smoothed_config_setup,
                    survey_windows, window_setup)

__all__ = ["EXPORTED", "LIB_PATH", "NTILES", "Qp3dError", "load", "Batch", "Fwi", "Plan", "modeling3d", "Setup", "build_setup",
           "config_setup", "fd_tables", "file_to_grid", "genmodel", "smoothed_config_setup", "survey_windows", "window_setup"]
