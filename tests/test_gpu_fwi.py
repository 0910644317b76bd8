"""Adjoint-state FWI gradient on the GPU: directional derivatives against central finite
differences of the misfit computed by the same GPU forward, checkpointed against
full-wavefield storage, BF16 checkpoints, and the torch.autograd shim.

The misfit is J(m) = 1/2 sum (record(m) - observed)^2 over all steps and receivers, with the
PML profiles held fixed. The finite differences are central differences of the misfit computed by
the FP64 CPU oracle, with model steps of 1e-4 of max(vp) or 1e-4 in epsilon and delta."""
import dataclasses
import os
import sys

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

import qp3d
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_setup(elliptic, nt=150, seed=4):
    s = qp3d.config_setup(hx=26, hy=24, nz0=22, pml=10, nt=nt, seed=seed, src_depth=3)
    if elliptic:
        s = dataclasses.replace(s, dele=s.eps.copy())
    return s


def observed(s):
    """Data of a model 4% slower with a smooth +5% anomaly below the source: the residual is
    of the size of the data, so the misfit's first derivative dominates its third in the
    central differences (with a residual far below the data the h^2 error term is relatively
    large)."""
    vp = s.vp.copy()
    z, x, y = np.meshgrid(np.arange(s.nz), np.arange(s.nx), np.arange(s.ny), indexing="ij")
    blob = np.exp(-(((z - s.nz * 0.6) / 5) ** 2 + ((x - s.nx / 2) / 6) ** 2 + ((y - s.ny / 2) / 6) ** 2))
    vp *= (0.96 * (1 + 0.05 * blob)).astype(np.float32)
    with qp3d.Plan(dataclasses.replace(s, vp=vp)) as p:
        return p.forward().astype(np.float64)


def misfit64(s, obs, bits=None):
    """J from the FP64 CPU oracle (an independent forward implementation of the same
    discrete operator), so that the central differences are free of FP32 rounding; with
    `bits`, every cell's threshold decision is replayed from a base run (frozen decisions)."""
    rec = O.modeling3d(s, "f64") if bits is None else O.record_frozen(s, bits)
    return 0.5 * np.sum((rec.astype(np.float64) - obs) ** 2)


def base_decisions(s):
    """The per-cell threshold decisions of the FP64 oracle's run at s, recorded for replay."""
    o = O.OracleState(s, "f64")
    try:
        bits = O.decision_bits(s)
        o.set_decisions(bits, 1)
        o.run_record()
    finally:
        o.close()
    return bits


def direction(s, seed):
    rng = np.random.default_rng(seed)
    d = gaussian_filter(rng.standard_normal((s.nz, s.nx, s.ny)), 2.0)
    return (d / np.abs(d).max()).astype(np.float32)


def fd_check(s, obs, which, seeds, rel_step, bits=None):
    """Directional derivatives <g, d> from the GPU adjoint and from FP64 central differences.
    The perturbed models are float32 arrays: the direction used on the adjoint side is the
    difference actually applied, (m+ - m-)/(2h)."""
    with qp3d.Plan(s) as p, qp3d.Fwi(p, interval=40) as f:
        rec = f.forward().astype(np.float64)
        g = f.gradient((rec - obs).astype(np.float32))
    names = ("vp", "eps", "dele")
    gi = g[names.index(which)]
    res = []
    for sd in seeds:
        base = getattr(s, which)
        d = direction(s, sd)
        h = rel_step * (float(np.abs(base).max()) if which == "vp" else 1.0)
        mp = (base.astype(np.float64) + h * d).astype(np.float32)
        mm = (base.astype(np.float64) - h * d).astype(np.float32)
        jp = misfit64(dataclasses.replace(s, **{which: mp}), obs, bits)
        jm = misfit64(dataclasses.replace(s, **{which: mm}), obs, bits)
        fd = (jp - jm) / (2 * h)
        ad = float(np.sum(gi * ((mp.astype(np.float64) - mm.astype(np.float64)) / (2 * h))))
        res.append((ad, fd, abs(ad - fd) / abs(fd)))
        say(f"{which} seed {sd}: adjoint {ad:.8e} fd {fd:.8e} rel {res[-1][2]:.2e}")
    return res


@pytest.mark.parametrize("which", ["vp", "eps", "dele"])
def test_gradient_matches_finite_differences_elliptic(which):
    """epsilon = delta: S = 1 on both sides of the threshold, so the misfit is smooth."""
    s = small_setup(elliptic=True)
    obs = observed(s)
    res = fd_check(s, obs, which, seeds=(1, 2, 3), rel_step=1e-4)
    assert max(r[2] for r in res) <= 1e-3


# Layered VTI model (epsilon > delta). Above the 0.01*G threshold the scalar operator depends
# on the gradient direction, below it S = 1: the free misfit jumps where a cell crosses the
# threshold, and a finite difference picks up every jump inside its step (with steps of 1e-5 a
# velocity direction measured 8e-4 and 1.3e-3). The differences here replay the base run's
# threshold decisions (frozen), which is the linearisation the adjoint transposes.
@pytest.mark.parametrize("which", ["vp", "eps", "dele"])
def test_gradient_matches_frozen_decision_differences_anisotropic(which):
    s = small_setup(elliptic=False)
    obs = observed(s)
    res = fd_check(s, obs, which, seeds=(1, 2), rel_step=1e-4, bits=base_decisions(s))
    assert max(r[2] for r in res) <= 1e-3


def test_checkpointed_equals_full_storage_and_bf16_checkpoints():
    s = small_setup(elliptic=False, nt=120)
    obs = observed(s)
    grads = {}
    for key, interval, ck in (("full", 120, "fp32"), ("ck7", 7, "fp32"), ("ck25", 25, "fp32"), ("bf16", 25, "bf16")):
        with qp3d.Plan(s) as p, qp3d.Fwi(p, interval=interval, ckpt_storage=ck) as f:
            rec = f.forward().astype(np.float64)
            grads[key] = f.gradient((rec - obs).astype(np.float32))
            say(f"{key}: {f.times()}")
    for key in ("ck7", "ck25", "bf16"):
        for a, b in zip(grads[key], grads["full"]):
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            say(f"{key} vs full: {err:.3e}")
            assert err <= (1e-5 if key != "bf16" else 2e-2)


@pytest.mark.parametrize("rec_plane", ["source depth", "mid depth"])
def test_rest_map_sweep_equals_full_sweep(rec_plane):
    """The adjoint sweep that leaves out the planes its touched map shows at rest gives the
    gradient of the sweep over every plane, value for value (zeros of either sign compare
    equal). The receiver plane at mid depth makes the adjoint wave spread up and down from it;
    pml 10 puts both y faces in folded tiles."""
    s = small_setup(elliptic=False, nt=120)
    if rec_plane == "mid depth":
        s = dataclasses.replace(s, sz=s.nz // 2)
    obs = observed(s)
    grads = {}
    for on in (True, False):
        with qp3d.Plan(s) as p, qp3d.Fwi(p, interval=25) as f:
            f.set_rest_map(on)
            rec = f.forward().astype(np.float64)
            # the second gradient of a run starts from the first's product fields and map (and
            # recomputes the last segment, which the first took from the forward run)
            g1 = f.gradient((rec - obs).astype(np.float32))
            grads[on] = f.gradient((rec - obs).astype(np.float32))
            for a, b in zip(g1, grads[on]):
                assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)
            say(f"rest map {on}: {f.times()}")
    for a, b in zip(grads[True], grads[False]):
        assert np.all(np.isfinite(a)) and np.abs(b).max() > 0
        assert np.array_equal(a, b), float(np.abs(a - b).max())


def test_resident_segments_equal_recomputed_ones():
    """Segments whose u_k stay resident after the forward run give the gradient of recomputed
    ones, on the first and on a second gradient (which recomputes the resident segments that a
    recomputation overwrote); changing the count asks for a new forward run."""
    s = small_setup(elliptic=False, nt=120)
    obs = observed(s)
    grads = {}
    with qp3d.Plan(s) as p, qp3d.Fwi(p, interval=25) as f:  # 5 segments
        for n in (1, 3, 5, -1):
            kept = f.set_resident_segments(n)
            assert kept == (5 if n < 0 else n)
            with pytest.raises(qp3d.Qp3dError):
                f.gradient(np.zeros((s.nt, s.hx, s.hy), np.float32))
            rec = f.forward().astype(np.float64)
            resid = (rec - obs).astype(np.float32)
            grads[n] = [f.gradient(resid), f.gradient(resid)]
            say(f"resident {kept}: {f.times()}")
        with pytest.raises(qp3d.Qp3dError):
            f.set_resident_segments(6)
    ref = grads[1][0]
    for n, gs in grads.items():
        for g in gs:
            for a, b in zip(g, ref):
                assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b), (n, float(np.abs(a - b).max()))


def test_predicted_threshold_gradient():
    """With gmax_mode 2 (predicted threshold and fix-up) the checkpoint recomputation
    reproduces the forward run's steps, so the checkpointed gradient matches full storage to
    the bar above, and the gradient matches gmax_mode 1's to rounding."""
    s = small_setup(elliptic=False, nt=120)
    obs = observed(s)
    grads = {}
    for key, mode, interval in (("g1", 1, 7), ("full", 2, 120), ("ck7", 2, 7)):
        with qp3d.Plan(s, gmax_mode=mode) as p, qp3d.Fwi(p, interval=interval) as f:
            rec = f.forward().astype(np.float64)
            grads[key] = f.gradient((rec - obs).astype(np.float32))
    for key, tol in (("ck7", 1e-5), ("g1", 1e-4)):
        for a, b in zip(grads[key], grads["full"]):
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            say(f"gmax_mode 2: {key} vs full storage {err:.3e}")
            assert err <= tol


@pytest.mark.parametrize("resident", [1, -1])
def test_autograd_shim_matches_direct_gradient(resident):
    """The shim with host tensors against the Fwi object's gradient (with one segment of u_k
    kept after the forward pass, and with all of them)."""
    import torch

    import qp3d.autograd

    s = small_setup(elliptic=False, nt=100)
    obs = observed(s)
    eng = qp3d.autograd.Engine(s, interval=30, resident=resident)
    try:
        vp = torch.tensor(s.vp, requires_grad=True)
        eps = torch.tensor(s.eps, requires_grad=True)
        de = torch.tensor(s.dele, requires_grad=True)
        rec = qp3d.autograd.modeling(vp, eps, de, eng)
        loss = 0.5 * ((rec.double() - torch.from_numpy(obs)) ** 2).sum()
        loss.backward()
    finally:
        eng.close()
    with qp3d.Plan(s) as p, qp3d.Fwi(p, interval=30) as f:
        r = f.forward().astype(np.float64)
        g = f.gradient((r - obs).astype(np.float32))
    np.testing.assert_array_equal(rec.detach().numpy(), r.astype(np.float32))
    for t, ref in zip((vp, eps, de), g):
        a = t.grad.numpy().astype(np.float64)
        assert np.linalg.norm(a - ref) <= 1e-6 * np.linalg.norm(ref)


def test_autograd_shim_with_gpu_tensors():
    """GPU tensors through the shim: the model goes to the C ABI by device pointer and the
    record and gradients come back as tensors on the GPU, identical to the host path. Runs in a
    child process that imports torch before libqp3d.so loads, so that both use one HIP runtime
    (this process loaded the library on its own runtime first)."""
    import subprocess

    child = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autograd_child.py")
    say("child process: " + child)
    r = subprocess.run([sys.executable, "-u", child], capture_output=True, text=True, timeout=240)
    say(r.stdout)
    say(r.stderr[-3000:])
    assert r.returncode == 0
