"""Host logic of the work decomposition (no GPU): the items the plan launches cover every
updated cell exactly once, interior items hold only interior cells, PML items only cells of
one face family, and tiles respect their shapes (qp3d_work_items, include/qp3d.h)."""
import ctypes as C

import numpy as np
import pytest

from qp3d._lib import QP3D_ERR_ARG, Config, check, load

TX = {1: 16, 2: 16, 3: 8, 4: 8, 5: 12, 6: 32, 7: 16}  # rows of x per tile shape (8x2 16x1 4x2 8x1 12x1 16x2 4x4)
H = 5


def items(cfg, which, capacity):
    lib = load()
    n = lib.qp3d_work_items(C.byref(cfg), which, capacity, None, 0)
    check(0 if n >= 0 else n)
    out = (C.c_int * (4 * max(n, 1)))()
    assert lib.qp3d_work_items(C.byref(cfg), which, capacity, out, n) == n
    a = np.frombuffer(out, np.int32)[:4 * n].reshape(n, 4).astype(np.int64)
    x0, x1 = a[:, 0] & 0xffff, (a[:, 0] >> 16) & 0xffff
    fold = a[:, 1] < 0
    yw = a[:, 1] & 0xffffffff
    y0, y1 = yw & 0xffff, (yw >> 16) & 0x7fff
    return x0, x1, y0, y1, a[:, 2], a[:, 3], fold


def cover(grid, x0, x1, y0, y1, z0, z1, fold, ny, pml):
    for i in range(len(x0)):
        ys = [(y0[i], y1[i])] if not fold[i] else [(y0[i], y1[i]), (ny - pml, ny - pml + (y1[i] - y0[i]))]
        for a, b in ys:
            grid[z0[i]:z1[i], x0[i]:x1[i], a:b] += 1


CASES = [
    # nx, ny, nz, pml, tile_update, tile_grad, tile_pml, tile_pml_y, zchunk, capacity
    (41, 63, 37, 10, 0, 0, 0, 0, 0, 256),
    (37, 71, 29, 9, 5, 4, 2, 3, 11, 64),
    (30, 33, 28, 6, 6, 6, 6, 6, 0, 1024),
    (90, 100, 80, 32, 1, 3, 5, 1, 0, 512),
    (95, 97, 88, 36, 2, 1, 6, 4, 7, 300),
    (85, 77, 90, 13, 4, 2, 1, 2, 0, 100),
    (70, 75, 66, 12, 7, 7, 7, 7, 9, 200),
    # ny % 4 == 0 with the default gradient tile: four y cells per lane, 64-aligned y segments
    (41, 64, 37, 10, 0, 0, 0, 0, 0, 256),
    (53, 140, 31, 12, 0, 0, 0, 0, 7, 96),
]


@pytest.mark.parametrize("case", CASES)
def test_work_items_cover_each_updated_cell_once(case):
    nx, ny, nz, pml, tu, tg, tp, tpy, zc, cap = case
    cfg = Config(nx, ny, nz, pml, 10, 12.5, 12.5, 12.5, 0.001, pml, 0, 0, 0, 0, zc, 0, tu, tg, tp, 0, tpy)
    upd = np.zeros((nz, nx, ny), np.int32)
    upd[H:nz - H, H:nx - H, H:ny - H] = 1
    interior = np.zeros_like(upd)
    interior[pml:nz - pml, pml:nx - pml, pml:ny - pml] = 1
    # gradient-max pass: uniform tiles
    g = np.zeros_like(upd)
    x0, x1, y0, y1, z0, z1, fold = items(cfg, 0, cap)
    assert not fold.any()
    assert np.all(x1 - x0 <= TX[tg or 7]) and np.all(y1 - y0 <= 64) and np.all(z1 > z0)
    cover(g, x0, x1, y0, y1, z0, z1, fold, ny, pml)
    assert np.array_equal(g, upd)
    # update: interior + PML x/z faces + PML y faces
    u = np.zeros_like(upd)
    x0, x1, y0, y1, z0, z1, fold = items(cfg, 1, cap)
    assert np.all(x1 - x0 <= TX[tu or 1])
    only = np.zeros_like(upd)
    cover(only, x0, x1, y0, y1, z0, z1, fold, ny, pml)
    assert not np.any(only & (1 - interior))
    u += only
    for which, t in ((2, tp or 6), (3, tpy or 2)):
        x0, x1, y0, y1, z0, z1, fold = items(cfg, which, cap)
        assert np.all(x1 - x0 <= TX[t])
        for i in range(len(x0)):
            c = np.zeros_like(upd)
            cover(c, x0[i:i + 1], x1[i:i + 1], y0[i:i + 1], y1[i:i + 1], z0[i:i + 1], z1[i:i + 1], fold[i:i + 1],
                  ny, pml)
            assert not np.any(c & interior)
            # one face family per item: x faces own edges and corners, then y, then z
            xf = x1[i] <= pml or x0[i] >= nx - pml
            yf = not xf and (fold[i] or y1[i] <= pml or y0[i] >= ny - pml)
            if which == 3:
                assert yf and not xf
            else:
                assert xf or (not yf and (z1[i] <= pml or z0[i] >= nz - pml))
            if fold[i]:
                assert y0[i] == H and y1[i] == pml and pml <= 32
            u += c
    assert np.array_equal(u, upd)


def test_work_items_rejects_bad_arguments():
    lib = load()
    cfg = Config(41, 41, 41, 4, 10, 12.5, 12.5, 12.5, 0.001, 10, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
    assert lib.qp3d_work_items(C.byref(cfg), 1, 256, None, 0) == QP3D_ERR_ARG  # pml < 6
    cfg.pml = 10
    assert lib.qp3d_work_items(C.byref(cfg), 4, 256, None, 0) < 0
    assert lib.qp3d_work_items(C.byref(cfg), 1, 0, None, 0) < 0
