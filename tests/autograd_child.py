"""Child process of test_gpu_fwi.py's GPU-tensor test of the torch.autograd shim. torch is
imported before anything loads libqp3d.so, so that the library binds to torch's HIP runtime and
device pointers cross between them. Runs one shot with the model as GPU tensors and compares
the record and the gradients with the Fwi object's (host arrays) on the same engine; prints the
figures and exits 0 when every check holds."""
import dataclasses
import os
import sys

import torch  # noqa: I001 (first, see above)

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "gpu-c_3dqmodeling_amd")):
    sys.path.insert(0, p)

import qp3d.autograd  # noqa: E402
import qp3d  # noqa: E402


def main():
    rt = qp3d.autograd.hip_runtimes()
    print("HIP runtimes mapped:", rt, flush=True)
    s = qp3d.config_setup(hx=26, hy=24, nz0=22, pml=10, nt=100, seed=4, src_depth=3)
    vp_obs = (s.vp * np.float32(0.97)).astype(np.float32)
    with qp3d.Plan(dataclasses.replace(s, vp=vp_obs)) as p:
        obs = p.forward().astype(np.float64)
    eng = qp3d.autograd.Engine(s, interval=30)
    try:
        dev = "cuda:0"
        vp = torch.tensor(s.vp, requires_grad=True, device=dev)
        eps = torch.tensor(s.eps, requires_grad=True, device=dev)
        de = torch.tensor(s.dele, requires_grad=True, device=dev)
        rec = qp3d.autograd.modeling(vp, eps, de, eng)
        assert rec.device == vp.device, rec.device
        loss = 0.5 * ((rec.double() - torch.from_numpy(obs).to(dev)) ** 2).sum()
        loss.backward()
        grads = [t.grad for t in (vp, eps, de)]
        assert all(g.device == vp.device and g.dtype == torch.float32 for g in grads)
        # the same engine with host arrays: record and gradient through host buffers
        eng.set_model(s.vp, s.eps, s.dele)
        r = eng.fwi.forward().astype(np.float64)
        g = eng.fwi.gradient((r - obs).astype(np.float32))
    finally:
        eng.close()
    ok = np.array_equal(rec.detach().cpu().numpy(), r.astype(np.float32))
    print("record identical:", ok, flush=True)
    for name, a, ref in zip(("vp", "eps", "dele"), grads, g):
        err = np.linalg.norm(a.cpu().numpy().astype(np.float64) - ref) / np.linalg.norm(ref)
        print(f"gradient {name}: rel {err:.3e}", flush=True)
        ok = ok and err <= 1e-6
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
