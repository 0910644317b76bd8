"""torch.autograd shim over the C ABI: records as a differentiable function of the padded-grid
vp, epsilon and delta. The forward pass runs a checkpointed shot on the GPU (qp3d_fwi_forward),
the backward pass back-propagates d(loss)/d(record) with the adjoint kernels
(qp3d_fwi_gradient). Torch is plumbing here. GPU tensors go to the C ABI by their device
pointers: the model is copied device to device into the plan, and the record and the gradients
are written straight into tensors on the plan's device. Host tensors go through host buffers.

    import qp3d.autograd                                  # before the first other qp3d call
    eng = qp3d.autograd.Engine(setup, interval=50)
    rec = qp3d.autograd.modeling(vp, eps, dele, eng)      # vp etc. require_grad
    loss = 0.5 * ((rec - observed) ** 2).sum(); loss.backward()

Device pointers cross between torch and libqp3d.so only when both run on one HIP runtime. A
torch wheel may carry its own HIP runtime (libamdhip64 with the system's soname); the dynamic
linker then binds libqp3d.so to torch's copy if torch is loaded first, which importing this
module before anything else loads libqp3d.so ensures. When the library was loaded first the
process holds two runtimes that do not know each other's allocations: GPU tensors are then
refused with an error, host tensors still work.
"""
import ctypes as C

import numpy as np
import torch

from ._lib import check, fptr
from .plan import Fwi, Plan, _tables


def _ptr(a, ctype=C.c_float):
    """Typed pointer of a contiguous float32/float64 GPU tensor or host array for the C ABI."""
    if isinstance(a, torch.Tensor):
        return C.cast(C.c_void_p(a.data_ptr()), C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


def _size(a):
    return a.numel() if isinstance(a, torch.Tensor) else a.size


def hip_runtimes():
    """Files of the HIP runtimes (libamdhip64) mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            return sorted({ln.split()[-1] for ln in f if "libamdhip64" in ln and "/" in ln})
    except OSError:
        return []


class Engine:
    """A plan and its gradient object for one grid, source and set of PML profiles; `resident`
    segments of u_k stay in HBM after the forward pass (qp3d_fwi_set_resident_segments; -1: as
    many as fit), which the backward pass does not recompute."""

    def __init__(self, setup, interval=50, ckpt_storage="fp32", device=0, resident=1):
        self.s = setup
        self.device = int(device)
        self.plan = Plan(setup, device=self.device)
        self.fwi = Fwi(self.plan, interval=interval, ckpt_storage=ckpt_storage)
        if resident != 1:
            self.fwi.set_resident_segments(resident)
        self._hold = []

    def _on_gpu(self, t):
        return isinstance(t, torch.Tensor) and t.is_cuda

    def _arg(self, t):
        """t as the C ABI takes it: a contiguous float32 tensor on the plan's device for a GPU
        tensor, a contiguous float32 host array otherwise."""
        if self._on_gpu(t):
            self._check_runtime()
            return t.detach().to(device=f"cuda:{self.device}", dtype=torch.float32).contiguous()
        if isinstance(t, torch.Tensor):
            t = t.detach().numpy()
        return np.ascontiguousarray(t, np.float32)

    def _check_runtime(self):
        rt = hip_runtimes()
        if len(rt) > 1:
            raise RuntimeError("GPU tensors need torch and libqp3d.so on one HIP runtime, but this process holds "
                               f"{len(rt)} ({', '.join(rt)}): import torch (or qp3d.autograd) before anything "
                               "loads libqp3d.so; host tensors work either way")

    def _fence(self):
        """The plan runs on a stream of its own: wait for the torch work that produced the
        tensors handed over."""
        torch.cuda.current_stream(self.device).synchronize()

    def set_model(self, vp, eps, dele):
        """Uploads vp, epsilon and delta ([nz][nx][ny]); source, wavelet and PML tables come
        from the setup. The arrays stay referenced until the next upload (the copies are
        queued on the plan's stream)."""
        model = [self._arg(t) for t in (vp, eps, dele)]
        rest = [np.ascontiguousarray(a, np.float32).reshape(-1) for a in [self.s.source, self.s.wavelet] + _tables(self.s)]
        n = self.s.nz * self.s.nx * self.s.ny
        if any(_size(m) != n for m in model):
            raise ValueError("vp, eps and dele must each hold nz*nx*ny values")
        if any(isinstance(m, torch.Tensor) for m in model):
            self._fence()
        check(self.plan.lib.qp3d_plan_upload(self.plan.h, *[_ptr(m) for m in model], *[fptr(a) for a in rest]))
        self._hold = model + rest

    def record(self, like):
        """Record [nt][hx][hy] of the last forward run, on like's device (float32)."""
        s = self.s
        if self._on_gpu(like):
            self._check_runtime()
            out = torch.empty((s.nt, s.hx, s.hy), dtype=torch.float32, device=f"cuda:{self.device}")
            self._fence()
            check(self.plan.lib.qp3d_plan_record(self.plan.h, _ptr(out)))
            return out.to(like.device)
        return torch.from_numpy(self.plan.record())

    def gradient(self, resid, like):
        """(g_vp, g_eps, g_del) [nz][nx][ny] float64 for resid = d(loss)/d(record), on like's
        device."""
        s = self.s
        r = self._arg(resid)
        if _size(r) != s.nt * s.hx * s.hy:
            raise ValueError("the record gradient must be [nt][hx][hy]")
        if self._on_gpu(like):
            out = [torch.empty((s.nz, s.nx, s.ny), dtype=torch.float64, device=f"cuda:{self.device}")
                   for _ in range(3)]
        else:
            out = [np.zeros((s.nz, s.nx, s.ny), np.float64) for _ in range(3)]
        if isinstance(r, torch.Tensor) or self._on_gpu(like):
            self._fence()
        check(self.fwi.lib.qp3d_fwi_gradient(self.fwi.h, _ptr(r), *[_ptr(o, C.c_double) for o in out]))
        return tuple(o if isinstance(o, torch.Tensor) else torch.from_numpy(o) for o in out)

    def close(self):
        self.fwi.close()
        self.plan.close()
        self._hold = []


class Modeling(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, eps, dele, engine):
        engine.set_model(vp, eps, dele)
        check(engine.fwi.lib.qp3d_fwi_forward(engine.fwi.h))
        ctx.engine = engine
        ctx.meta = [(t.dtype, t.device) for t in (vp, eps, dele)]
        return engine.record(vp)

    @staticmethod
    def backward(ctx, grad_rec):
        g = ctx.engine.gradient(grad_rec, grad_rec)
        out = tuple(a.to(device=dev, dtype=dt) for a, (dt, dev) in zip(g, ctx.meta))
        return out + (None,)


def modeling(vp, eps, dele, engine):
    """Record [nt][hx][hy] of one shot as a differentiable function of the model."""
    return Modeling.apply(vp, eps, dele, engine)
