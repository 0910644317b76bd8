"""Per-shot setup in the order of the reference driver (Main.c:125-201), on top of the
library's host setup functions (bit-identical to CPU_function.c's tables)."""
from dataclasses import dataclass, field

import numpy as np

from ._lib import fptr, load, check


@dataclass
class Setup:
    """Everything modeling3d() takes for one shot (3dPureqPModeling_single_shot.h:4-6)."""
    nx: int
    ny: int
    nz: int
    pml: int
    nt: int
    dx: float
    dy: float
    dz: float
    dt: float
    sz: int                 # receiver plane (padded z index)
    vp: np.ndarray          # [nz][nx][ny] float32, padded
    eps: np.ndarray
    dele: np.ndarray
    source: np.ndarray
    wavelet: np.ndarray     # [nt]
    dlr: np.ndarray         # [pml] each
    ddlr: np.ndarray
    dtb: np.ndarray
    ddtb: np.ndarray
    dfb: np.ndarray
    ddfb: np.ndarray
    c: np.ndarray           # [36]
    c2: np.ndarray          # [36]
    src_cell: tuple = field(default=(0, 0, 0))  # (sx, sy, sz) of the source centre

    @property
    def hx(self):
        return self.nx - 2 * self.pml

    @property
    def hy(self):
        return self.ny - 2 * self.pml

    @property
    def hz(self):
        return self.nz - 2 * self.pml

    @property
    def n_upd(self):
        return (self.nx - 10) * (self.ny - 10) * (self.nz - 10)

    @property
    def n_int(self):
        return self.hx * self.hy * self.hz


def fd_tables():
    lib = load()
    c = np.zeros(36, np.float32)
    c2 = np.zeros(36, np.float32)
    lib.qp3d_fdcoff1(fptr(c))
    lib.qp3d_fdcoff2(fptr(c2))
    return c, c2


def genmodel(hx, hy, nz0, seed, sigma=5.0, amp=0.03):
    """Layered VTI model + smooth random perturbation, returned in model-file layout
    [hy][hx][nz0] (what readmodel3d reads) as (vp, eps, del)."""
    lib = load()
    n = hx * hy * nz0
    vp = np.empty(n, np.float32)
    ep = np.empty(n, np.float32)
    de = np.empty(n, np.float32)
    check(lib.qp3d_genmodel(hx, hy, nz0, int(seed), float(sigma), float(amp), fptr(vp), fptr(ep), fptr(de)))
    shape = (hy, hx, nz0)
    return vp.reshape(shape), ep.reshape(shape), de.reshape(shape)


def file_to_grid(a):
    """Model-file layout [y][x][z] -> solver layout [z][x][y] (CPU_function.c:254-259)."""
    return np.ascontiguousarray(np.transpose(a, (2, 1, 0)))


def build_setup(vp1, eps1, del1, pml, nt, dx, dy, dz, dt, fm, amp=1.0, alpha=1.0, src_depth=0, rec_depth=None):
    """vp1/eps1/del1: interior model [nz0][hx][hy] float32 (solver layout).
    Mirrors Main.c:191-201: extend_model, vpmax, pmlcoff, source at
    ((hx-1)/2+pml, (hy-1)/2+pml, pml+src_depth), Ricker with t0 = 1/fm.
    src_depth = 0 reproduces the reference (Main.c:200 forces sz = pml); rec_depth defaults
    to src_depth (the record plane is modeling3d's sz, which Main.c sets to the source depth)."""
    lib = load()
    nz0, hx, hy = vp1.shape
    nx, ny, nz = hx + 2 * pml, hy + 2 * pml, nz0 + 2 * pml
    grid = []
    for a in (vp1, eps1, del1):
        a = np.ascontiguousarray(a, np.float32)
        out = np.zeros(nx * ny * nz, np.float32)
        lib.qp3d_extend_model(fptr(a.reshape(-1)), fptr(out), nx, ny, nz, pml)
        grid.append(out.reshape(nz, nx, ny))
    vp, eps, dele = grid
    n = nx * ny * nz
    f32 = np.float32
    vpmax = (f32(1) + f32(2) * f32(lib.qp3d_absmax(fptr(eps.reshape(-1)), n))) * f32(lib.qp3d_absmax(fptr(vp.reshape(-1)), n))
    tabs = [np.zeros(pml, np.float32) for _ in range(6)]
    lib.qp3d_pmlcoff(pml, float(vpmax), dx, dy, dz, *[fptr(t) for t in tabs])
    sx1 = (hx - 1) // 2 + pml
    sy1 = (hy - 1) // 2 + pml
    sz1 = pml + src_depth
    t0 = float(np.float32(1.0 / fm))
    source = np.zeros(n, np.float32)
    wavelet = np.zeros(nt, np.float32)
    lib.qp3d_source(sx1, sy1, sz1, fm, amp, alpha, dt, dx, dy, dz, t0, nt, nx, ny, nz, fptr(source), fptr(wavelet))
    c, c2 = fd_tables()
    rz = pml + (src_depth if rec_depth is None else rec_depth)
    return Setup(nx, ny, nz, pml, nt, float(dx), float(dy), float(dz), float(dt), rz, vp, eps, dele,
                 source.reshape(nz, nx, ny), wavelet, *tabs, c, c2, src_cell=(sx1, sy1, sz1))


def config_setup(hx, hy, nz0, pml, nt, seed, h=12.5, dt=0.001, fm=25.0, src_depth=0, sigma=5.0, amp=0.03):
    """A shot of the synthetic layered-plus-smooth-random VTI model rule."""
    vpf, epf, def_ = genmodel(hx, hy, nz0, seed, sigma, amp)
    return build_setup(file_to_grid(vpf), file_to_grid(epf), file_to_grid(def_), pml, nt, h, h, h, dt, fm,
                       src_depth=src_depth)


def window_setup(model, x0, y0, hx, hy, pml, nt, fm, h=12.5, dt=0.001, src_depth=0):
    """Setup of one shot of a survey (Main.c:179-201): the window [x0, x0+hx) x [y0, y0+hy) x
    [0, nz0) of a model held in model-file layout [y][x][z] (as readmodel3d reads it), with the
    source at the window centre."""
    win = [np.ascontiguousarray(a[y0:y0 + hy, x0:x0 + hx, :]) for a in model]
    if any(w.shape[:2] != (hy, hx) for w in win):
        raise ValueError("window outside the model")
    return build_setup(*[file_to_grid(w) for w in win], pml, nt, h, h, h, dt, fm, src_depth=src_depth)


def survey_windows(nx0, ny0, nshots, wmin, wmax, seed):
    """Shot windows of a survey over an nx0 x ny0 model surface: shot positions uniform on the
    surface grid, each window wmin..wmax cells in x and y (independently), centred on its shot
    and moved inside the model where it would cross an edge. Returns [(x0, y0, wx, wy)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nshots):
        wx, wy = (int(v) for v in rng.integers(wmin, wmax + 1, size=2))
        sx, sy = int(rng.integers(0, nx0)), int(rng.integers(0, ny0))
        x0 = min(max(sx - wx // 2, 0), nx0 - wx)
        y0 = min(max(sy - wy // 2, 0), ny0 - wy)
        out.append((x0, y0, wx, wy))
    return out


def smoothed_config_setup(hx, hy, nz0, pml, nt, seed, sigma, fm=25.0, src_depth=0, h=12.5, dt=0.001):
    """The shot of config_setup with its model smoothed by a Gaussian of `sigma` cells (the
    starting model of an inversion built from the same model rule)."""
    from scipy.ndimage import gaussian_filter
    vp, ep, de = genmodel(hx, hy, nz0, seed)
    sm = [gaussian_filter(a.astype(np.float64), sigma, mode="nearest").astype(np.float32) for a in (vp, ep, de)]
    return build_setup(*[file_to_grid(a) for a in sm], pml, nt, h, h, h, dt, fm, src_depth=src_depth)
