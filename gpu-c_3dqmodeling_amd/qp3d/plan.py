"""Python mirror of the solver boundary: modeling3d() (3dPureqPModeling_single_shot.h:4-6)
and the device-resident plan API, both through the C ABI of libqp3d.so."""
import ctypes as C

import numpy as np

from ._lib import Config, STORAGE, check, fptr, load


def _tables(s):
    return [s.dlr, s.ddlr, s.dtb, s.ddtb, s.dfb, s.ddfb, s.c, s.c2]


def modeling3d(s, record=None, snapflag=0, ntsnap=1, snap_file=None):
    """Run one shot through the drop-in entry point qp3d_modeling3d (same arguments as the
    reference's modeling3d). Returns the record [nt][hx][hy]."""
    lib = load()
    if record is None:
        record = np.zeros((s.nt, s.hx, s.hy), np.float32)
    sx, sy, _ = s.src_cell
    arrs = [s.vp, s.eps, s.dele, s.source, s.wavelet, record] + _tables(s)
    arrs = [np.ascontiguousarray(a, np.float32) for a in arrs]
    rc = lib.qp3d_modeling3d(s.nx, s.ny, s.nz, s.nt, ntsnap, s.dx, s.dy, s.dz, s.dt, s.pml, snapflag, sx, sy, s.sz,
                             *[fptr(a.reshape(-1)) for a in arrs],
                             snap_file.encode() if snap_file else None)
    check(rc)
    return arrs[5].reshape(s.nt, s.hx, s.hy)


class Plan:
    """Device-resident forward solver for one grid (qp3d_plan_* in include/qp3d.h)."""

    def __init__(self, s, storage="fp32", use_graph=False, snapflag=0, ntsnap=0, zchunk=0, device=0,
                 tile_update=0, tile_grad=0, tile_pml=0, streams=0, tile_pml_y=0, snap_s=0, gmax_mode=0):
        self.lib = load()
        self.s = s
        cfg = Config(s.nx, s.ny, s.nz, s.pml, s.nt, s.dx, s.dy, s.dz, s.dt, s.sz, STORAGE[storage],
                     int(use_graph), int(snapflag), int(ntsnap), int(zchunk), int(device),
                     int(tile_update), int(tile_grad), int(tile_pml), int(streams), int(tile_pml_y),
                     int(snap_s), int(gmax_mode))
        h = C.c_void_p()
        check(self.lib.qp3d_plan_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.upload(s)

    def upload(self, s):
        arrs = [np.ascontiguousarray(a, np.float32).reshape(-1)
                for a in [s.vp, s.eps, s.dele, s.source, s.wavelet] + _tables(s)]
        check(self.lib.qp3d_plan_upload(self.h, *[fptr(a) for a in arrs]))
        self.s = s

    def reset(self):
        check(self.lib.qp3d_plan_reset(self.h))

    def run(self, k0=0, k1=None):
        check(self.lib.qp3d_plan_run(self.h, k0, self.s.nt if k1 is None else k1))

    def sync(self):
        check(self.lib.qp3d_plan_sync(self.h))

    def forward_async(self):
        """Enqueue one full shot from rest (reset + all time steps) without waiting."""
        self.reset()
        self.run()

    def forward(self):
        self.forward_async()
        self.sync()
        return self.record()

    def record(self):
        s = self.s
        out = np.zeros((s.nt, s.hx, s.hy), np.float32)
        check(self.lib.qp3d_plan_record(self.h, fptr(out.reshape(-1))))
        return out

    def set_amp_bound(self, on=True):
        """Amplitude bound of the gradient pass on or off (results are the same)."""
        check(self.lib.qp3d_plan_set_amp_bound(self.h, int(bool(on))))

    def gradient_tiles(self):
        """(evaluated, skipped) gradient-pass tiles since the last reset."""
        out = (C.c_ulonglong * 2)()
        check(self.lib.qp3d_plan_gradient_tiles(self.h, out))
        return int(out[0]), int(out[1])

    def update_cells(self):
        """Per update kind (interior, PML over the full grid, folded y faces): (cells of its
        items, left out at rest, updated without eps - del) since the last reset, counted on
        instrumented steps; a 3 x 3 integer array."""
        out = (C.c_ulonglong * 9)()
        check(self.lib.qp3d_plan_update_cells(self.h, out))
        return np.array(out[:], dtype=np.int64).reshape(3, 3)

    def field(self, which=0):
        s = self.s
        out = np.zeros((s.nz, s.nx, s.ny), np.float32)
        check(self.lib.qp3d_plan_field(self.h, int(which), fptr(out.reshape(-1))))
        return out

    def snapshots(self):
        s = self.s
        res = []
        for i in range(self.lib.qp3d_plan_snapshot_count(self.h)):
            out = np.zeros((s.hz, s.hx, s.hy), np.float32)
            k = C.c_int()
            check(self.lib.qp3d_plan_snapshot(self.h, i, C.byref(k), fptr(out.reshape(-1))))
            res.append((k.value, out))
        return res

    def snapshots_s(self):
        """[(step, S)] of the snapshot steps, S over the padded grid [nz][nx][ny] (snap_s=1)."""
        s = self.s
        res = []
        for i in range(self.lib.qp3d_plan_snapshot_count(self.h)):
            out = np.zeros((s.nz, s.nx, s.ny), np.float32)
            k = C.c_int()
            check(self.lib.qp3d_plan_snapshot_s(self.h, i, C.byref(k), fptr(out.reshape(-1))))
            res.append((k.value, out))
        return res

    def step_maxima(self):
        out = np.zeros(self.s.nt, np.float32)
        check(self.lib.qp3d_plan_step_maxima(self.h, fptr(out)))
        return out

    def step_predictions(self):
        """gmax_mode 2: the predicted G of every step (the threshold and band it was run with)."""
        out = np.zeros(self.s.nt, np.float32)
        check(self.lib.qp3d_plan_step_predictions(self.h, fptr(out)))
        return out

    def gmax_mode(self):
        """1: a gradient pass gives G before each update; 2: predicted threshold with fix-up."""
        rc = self.lib.qp3d_plan_gmax_mode(self.h)
        check(min(rc, 0))
        return rc

    def set_instrument(self, on=True):
        """True / 1: HIP events around every stencil launch; 2: around the interior update
        only; False / 0: none. Resets the accumulated times."""
        check(self.lib.qp3d_plan_set_instrument(self.h, int(on)))

    def kernel_times(self):
        """Mean ms per step of the gradient-max, interior-update and PML-update launches, of the
        whole update stage (the two updates overlap on two streams), and the number of
        instrumented steps."""
        ms = (C.c_double * 5)()
        check(self.lib.qp3d_plan_kernel_times(self.h, ms))
        return {"gradmax": ms[0], "update_interior": ms[1], "update_pml": ms[2], "update_stage": ms[3],
                "steps": int(ms[4])}

    def step_times(self):
        """Interior-update time (ms) of each step from the last instrumented run (0: not measured)."""
        out = np.zeros(self.s.nt, np.float32)
        check(self.lib.qp3d_plan_step_times(self.h, fptr(out)))
        return out

    def step_bytes(self):
        """Algorithmic bytes per step of each launch and the cells of the two update launches."""
        b = (C.c_double * 5)()
        check(self.lib.qp3d_plan_step_bytes(self.h, b))
        return {"gradmax": b[0], "update_interior": b[1], "update_pml": b[2], "cells_interior": b[3],
                "cells_pml": b[4]}

    def close(self):
        if self.h:
            self.lib.qp3d_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Fwi:
    """Adjoint-state gradient on top of a Plan (qp3d_fwi_* in include/qp3d.h): a checkpointed
    forward run, then the gradient of 1/2 |record - observed|^2 w.r.t. the padded-grid vp,
    epsilon and delta, by segment recomputation and the adjoint kernels."""

    def __init__(self, plan, interval=50, ckpt_storage="fp32"):
        self.plan = plan
        self.lib = plan.lib
        h = C.c_void_p()
        check(self.lib.qp3d_fwi_create(plan.h, int(interval), STORAGE[ckpt_storage], C.byref(h)))
        self.h = h

    def forward(self):
        """Runs the shot from rest with checkpoints and returns the record [nt][hx][hy]."""
        check(self.lib.qp3d_fwi_forward(self.h))
        return self.plan.record()

    def gradient(self, resid):
        """resid = record - observed, [nt][hx][hy]; returns (g_vp, g_eps, g_del) [nz][nx][ny] float64."""
        s = self.plan.s
        r = np.ascontiguousarray(resid, np.float32).reshape(-1)
        if r.size != s.nt * s.hx * s.hy:
            raise ValueError("residual must be [nt][hx][hy]")
        out = [np.zeros((s.nz, s.nx, s.ny), np.float64) for _ in range(3)]
        check(self.lib.qp3d_fwi_gradient(self.h, fptr(r), *[_dptr(o) for o in out]))
        return tuple(out)

    def set_rest_map(self, on):
        """Whether the adjoint sweep leaves out the planes its touched map shows at rest (default on)."""
        check(self.lib.qp3d_fwi_set_rest_map(self.h, 1 if on else 0))

    def set_resident_segments(self, n):
        """Keeps the u_k of the last n segments in HBM after the forward run (n < 0: as many as
        fit); returns the number kept. Run forward() again afterwards."""
        kept = C.c_int(0)
        check(self.lib.qp3d_fwi_set_resident_segments(self.h, int(n), C.byref(kept)))
        return kept.value

    def times(self):
        ms = (C.c_double * 3)()
        check(self.lib.qp3d_fwi_times(self.h, ms))
        return {"forward": ms[0], "recompute": ms[1], "adjoint": ms[2]}

    def close(self):
        if self.h:
            self.lib.qp3d_fwi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """Several shots (one Plan each) advanced in lockstep with one launch per stage and step
    (qp3d_batch_* in include/qp3d.h)."""

    def __init__(self, plans, use_graph=False):
        self.plans = list(plans)
        self.lib = self.plans[0].lib
        arr = (C.c_void_p * len(self.plans))(*[p.h.value for p in self.plans])
        h = C.c_void_p()
        check(self.lib.qp3d_batch_create(arr, len(self.plans), int(use_graph), C.byref(h)))
        self.h = h

    def reset(self):
        check(self.lib.qp3d_batch_reset(self.h))

    def run(self, k0=0, k1=None):
        check(self.lib.qp3d_batch_run(self.h, k0, self.plans[0].s.nt if k1 is None else k1))

    def sync(self):
        check(self.lib.qp3d_batch_sync(self.h))

    def forward(self):
        """All shots from rest; returns their records."""
        self.reset()
        self.run()
        self.sync()
        return [p.record() for p in self.plans]

    def autotune(self, steps=5):
        tiles = (C.c_int * 4)()
        ms = (C.c_double * 4)()
        check(self.lib.qp3d_batch_autotune(self.h, int(steps), tiles, ms))
        return {"tiles": list(tiles), "ms": list(ms), "per_shot": bool(self.lib.qp3d_batch_per_shot(self.h)),
                "share": bool(self.lib.qp3d_batch_share(self.h)), "persistent": self.persistent()}

    def persistent(self):
        """Tile set of the one-launch-per-run persistent mode, or -1 when it is off."""
        return int(self.lib.qp3d_batch_persistent(self.h))

    def set_persistent(self, tile_set=0, share=False):
        """Persistent mode with a tile set (0, 1, 2), or off (tile_set=-1)."""
        check(self.lib.qp3d_batch_set_persistent(self.h, int(tile_set), int(bool(share))))

    def persistent_choice(self):
        """(tile set, share) of the fastest persistent form of the last autotune."""
        a, b = C.c_int(), C.c_int()
        check(self.lib.qp3d_batch_persistent_choice(self.h, C.byref(a), C.byref(b)))
        return a.value, bool(b.value)

    def set_wait_limit(self, us):
        """Bound (microseconds) of one device-side wait of the persistent launch."""
        check(self.lib.qp3d_batch_set_wait_limit(self.h, float(us)))

    def set_per_shot(self, on=True):
        """Each shot's own launches within a step (True) or one launch per stage (False)."""
        check(self.lib.qp3d_batch_set_per_shot(self.h, int(bool(on))))

    def step_bytes(self):
        b = (C.c_double * 4)()
        check(self.lib.qp3d_batch_step_bytes(self.h, b))
        return {"gradmax": b[0], "update": b[1], "cells_updated": b[2], "cells_interior": b[3]}

    def close(self):
        if self.h:
            self.lib.qp3d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
