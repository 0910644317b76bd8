"""Amplitude bound of the gradient pass: the pass skips tiles whose bound 2 sum|c1|/d max|u|
(from the block maxima the updates record) is below the G already found exactly, so G_k and
every result must be the same with the bound on and off, bit for bit: direct launches and the
captured graph, 16-bit storage, a source on the frozen rim, and a plan whose block maxima went
stale while a batch drove it. The bound must also actually skip tiles on a shot from rest."""
import sys

import numpy as np
import pytest

import qp3d

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def state(p):
    return (p.record(), p.field(0), p.field(1), p.field(2), p.field(4), p.field(6), p.field(8), p.step_maxima())


def run(p, split=None):
    p.reset()
    if split:
        p.run(0, split)
        p.run(split, p.s.nt)
    else:
        p.run()
    p.sync()
    return state(p)


def same(a, b):
    for j, (u, v) in enumerate(zip(a, b)):
        assert np.array_equal(u, v), (j, float(np.abs(u - v).max()))


@pytest.mark.parametrize("graph", [False, True])
def test_bound_on_and_off_give_the_same_shot(graph):
    s = qp3d.config_setup(hx=101, hy=101, nz0=101, pml=20, nt=400, seed=0, src_depth=3)
    with qp3d.Plan(s, use_graph=graph) as p:
        p.set_amp_bound(False)
        say(f"bound off, graph={graph}")
        ref = run(p)
        assert p.gradient_tiles() == (0, 0)
        p.set_amp_bound(True)
        say(f"bound on, graph={graph}")
        got = run(p, split=37)
        same(got, ref)
        p.set_instrument(1)  # tiles are counted on instrumented steps
        same(run(p), ref)
        p.set_instrument(0)
        ev, sk = p.gradient_tiles()
        say(f"gradient tiles evaluated {ev}, skipped {sk}")
        same(got, ref)
        assert sk > 0.5 * (ev + sk)  # from rest most of the grid is far from the wavefront


@pytest.mark.parametrize("storage", ["bf16", "fp16"])
def test_bound_with_16bit_storage(storage):
    s = qp3d.config_setup(hx=61, hy=57, nz0=53, pml=12, nt=250, seed=1, src_depth=2)
    with qp3d.Plan(s, storage=storage) as p:
        p.set_amp_bound(False)
        ref = run(p)
        p.set_amp_bound(True)
        p.set_instrument(1)
        say(f"bound on, {storage}")
        same(run(p), ref)
        assert p.gradient_tiles()[1] > 0


def test_bound_with_a_source_on_the_frozen_rim():
    s = qp3d.config_setup(hx=41, hy=45, nz0=37, pml=10, nt=200, seed=2, src_depth=0)
    src = np.zeros_like(s.source)
    src[2, s.nx // 2, s.ny // 2] = 1.0       # rim (z < 5)
    src[s.nz // 2, 3, s.ny // 3] = -0.5      # rim (x < 5)
    src[s.nz // 2, s.nx // 2, s.ny // 2] = 0.25
    s.source = src
    with qp3d.Plan(s) as p:
        p.set_amp_bound(False)
        ref = run(p)
        p.set_amp_bound(True)
        say("bound on, rim source")
        same(run(p), ref)


def test_bound_after_a_batch_drove_the_plan():
    """A batch leaves the plan's block maxima stale: the plan's next run from the batch's state
    must evaluate every tile of its first step (no skipping) and match a run without the bound."""
    model = qp3d.genmodel(70, 64, 30, 3)
    setups = [qp3d.window_setup(model, x0, y0, hx, hy, pml=10, nt=60, fm=20.0, src_depth=2)
              for x0, y0, hx, hy in [(0, 0, 31, 27), (20, 9, 44, 40)]]
    plans = [qp3d.Plan(s) for s in setups]
    try:
        outs = []
        for bound in (False, True):
            for p in plans:
                p.set_amp_bound(bound)
            with qp3d.Batch(plans) as b:
                b.reset()
                b.run(0, 30)
                b.sync()
            say(f"plan runs on after the batch, bound {bound}")
            plans[1].run(30, 60)
            plans[1].sync()
            outs.append(state(plans[1]))
        same(outs[1], outs[0])
    finally:
        for p in plans:
            p.close()
