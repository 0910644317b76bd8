"""torch.autograd shim over the C ABI: records as a differentiable function of the padded-grid
vp, epsilon and delta. The forward pass runs a checkpointed shot on the GPU (qp3d_fwi_forward),
the backward pass back-propagates d(loss)/d(record) with the adjoint kernels
(qp3d_fwi_gradient). Torch is plumbing here: tensors are copied to host buffers for the C ABI.

    eng = qp3d.autograd.Engine(setup, interval=50)
    rec = qp3d.autograd.modeling(vp, eps, dele, eng)      # vp etc. require_grad
    loss = 0.5 * ((rec - observed) ** 2).sum(); loss.backward()
"""
import dataclasses

import numpy as np
import torch

from .plan import Fwi, Plan


class Engine:
    """A plan and its gradient object for one grid, source and set of PML profiles."""

    def __init__(self, setup, interval=50, ckpt_storage="fp32", device=0):
        self.s = setup
        self.plan = Plan(setup, device=device)
        self.fwi = Fwi(self.plan, interval=interval, ckpt_storage=ckpt_storage)

    def set_model(self, vp, eps, dele):
        self.s = dataclasses.replace(self.s, vp=vp, eps=eps, dele=dele)
        self.plan.upload(self.s)

    def close(self):
        self.fwi.close()
        self.plan.close()


def _host(t):
    return np.ascontiguousarray(t.detach().to("cpu", torch.float32).numpy())


class Modeling(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, eps, dele, engine):
        engine.set_model(_host(vp), _host(eps), _host(dele))
        rec = engine.fwi.forward()
        ctx.engine = engine
        ctx.meta = [(t.dtype, t.device) for t in (vp, eps, dele)]
        return torch.from_numpy(rec).to(vp.device)

    @staticmethod
    def backward(ctx, grad_rec):
        g = ctx.engine.fwi.gradient(_host(grad_rec))
        out = tuple(torch.from_numpy(a).to(device=dev, dtype=dt) for a, (dt, dev) in zip(g, ctx.meta))
        return out + (None,)


def modeling(vp, eps, dele, engine):
    """Record [nt][hx][hy] of one shot as a differentiable function of the model."""
    return Modeling.apply(vp, eps, dele, engine)
